"""Known-answer tests of the oracle, written by hand from the reference code.

Four parts of the oracle restate reference behaviour that no closed form or
second statement pins, so each gets expected values worked out by hand from
the reference source:

- SOAP indel edits (cpp/soap_snp.h:143-163): insertion padding with 'N' in
  read and quality, the deletion cut, and std::string::substr clipping at the
  read-length field;
- the calibration matrix fall-backs (cpp/matrix.cc:129-205): per cycle when a
  count exceeds 10, per base pair type, per mismatch rate, and the Phred
  values with their 0.25 bounds where a rate is 0 or 1;
- the .cns row forms (cpp/call_genotype.cc:355-406): the N row literal, the
  base2 >= 4 form, full rows with showpoint columns, integer-division
  average qualities;
- one -C window chunk of each column file, assembled byte by byte from
  cpp/compress.h:151-1312 and cpp/dict_compression.h:116-166.

A one-field swap or an off-by-one in any of these changes an expected value.
CPU only; the matrix case is also run through the product on the GPU
(tests/test_gpu_kats.py).
"""
from __future__ import annotations

import ctypes
import math
import os
import struct
from pathlib import Path

import numpy as np
import pytest

import oracle_lib


def _b(s) -> bytes:
    return os.fsencode(str(s))


# ------------------------------------------------------------------ SOAP edits
def _soap(i, seq, qual, pos1, mm, ip=None, ln=None, strand="+", hit=1, chrom="chr1"):
    ln = len(seq) if ln is None else ln
    tail = f"\t{ip}" if ip is not None else ""
    return f"r{i}\t{seq}\t{qual}\t{hit}\ta\t{ln}\t{strand}\t{chrom}\t{pos1}\t{mm}{tail}\n"


# (SOAP line fields, expected stored read and quality at pitch 12), each
# expected string worked out from cpp/soap_snp.h:143-163:
#   insertion (mismatch > 200): read.substr(0,ip) + 'N'*(mm-200) + read.substr(ip, read_len-ip)
#   deletion (mismatch > 100):  read.substr(0,ip) + read.substr(ip+il, read_len-ip-il)
# substr clips its count at the string end; past the edited string the oracle
# stores 'N' and quality 0 (the reference copies read_len bytes of c_str():
# the terminating NUL, then bytes it does not own)
SOAP_KATS = [
    # insertion of 2 at the start
    (("ACGTACGTAC", "ABCDEFGHIJ", 202, 0, None), "NNACGTACGTAC", "NNABCDEFGHIJ"),
    # insertion of 1 in the middle: one byte past the edit is padding
    (("ACGTACGTAC", "ABCDEFGHIJ", 201, 4, None), "ACGTNACGTACN", "ABCDNEFGHIJ\0"),
    # insertion of 2 at the end (substr(10, 0) is empty)
    (("ACGTACGTAC", "ABCDEFGHIJ", 202, 10, None), "ACGTACGTACNN", "ABCDEFGHIJNN"),
    # deletion of 2 at the start: substr(2, 8)
    (("ACGTACGTAC", "ABCDEFGHIJ", 102, 0, None), "GTACGTACNNNN", "CDEFGHIJ\0\0\0\0"),
    # deletion of 1 in the middle: "ACG" + substr(4, 6)
    (("ACGTACGTAC", "ABCDEFGHIJ", 101, 3, None), "ACGACGTACNNN", "ABCEFGHIJ\0\0\0"),
    # deletion of 2 at the end: substr(10, 0)
    (("ACGTACGTAC", "ABCDEFGHIJ", 102, 8, None), "ACGTACGTNNNN", "ABCDEFGH\0\0\0\0"),
    # read string longer than its length field (10): the tail after the
    # insertion is clipped to read_len - ip = 8 characters
    (("ACGTACGTACGG", "ABCDEFGHIJKL", 201, 2, 10), "ACNGTACGTACN", "ABNCDEFGHIJ\0"),
    # the same for a deletion: substr(5, 10 - 3 - 2 = 5)
    (("ACGTACGTACGG", "ABCDEFGHIJKL", 102, 3, 10), "ACGCGTACNNNN", "ABCFGHIJ\0\0\0\0"),
    # no indel (mismatch count 3): stored as read, padded past its length
    (("ACGTACGTAC", "ABCDEFGHIJ", 3, None, None), "ACGTACGTACNN", "ABCDEFGHIJ\0\0"),
]


def test_soap_indel_edits_by_hand(tmp_path):
    o = oracle_lib.orc()
    fa = tmp_path / "g.fa"
    fa.write_text(">chr1\n" + "ACGT" * 100 + "\n")
    lines = []
    for i, ((seq, qual, mm, ip, ln), _, _) in enumerate(SOAP_KATS):
        lines.append(_soap(i, seq, qual, 11 + 20 * i, mm, ip, ln))
    so = tmp_path / "r.soap"
    so.write_text("".join(lines))
    p = oracle_lib.params(read_length=12)
    g, r = ctypes.c_void_p(), ctypes.c_void_p()
    assert o.orc_genome_load(_b(fa), None, ctypes.byref(g)) == 0
    try:
        assert o.orc_reads_load(_b(so), ctypes.byref(p), g, ctypes.byref(r)) == 0, o.orc_last_error()
        rd = ctypes.cast(r, ctypes.POINTER(oracle_lib.OrcReads)).contents
        assert rd.n == len(SOAP_KATS) and rd.L == 12
        seq = ctypes.string_at(rd.seq, rd.n * 12)
        qual = ctypes.string_at(rd.qual, rd.n * 12)
        for i, ((s, _, _, _, ln), want_s, want_q) in enumerate(SOAP_KATS):
            assert rd.pos[i] == 11 + 20 * i - 1, "SOAP positions are 1-based"
            assert rd.len[i] == (ln if ln is not None else len(s)), "the length field is kept, not the edited length"
            assert seq[12 * i:12 * i + 12] == want_s.encode(), (i, seq[12 * i:12 * i + 12])
            assert qual[12 * i:12 * i + 12] == want_q.encode("latin-1"), (i, qual[12 * i:12 * i + 12])
        o.orc_reads_free(r)
    finally:
        o.orc_genome_free(g)


def test_soap_indel_position_outside_the_read_is_an_error(tmp_path):
    # the reference throws std::out_of_range from substr; the oracle reports a format error
    o = oracle_lib.orc()
    fa = tmp_path / "g.fa"
    fa.write_text(">chr1\n" + "ACGT" * 100 + "\n")
    so = tmp_path / "r.soap"
    so.write_text(_soap(0, "ACGTACGTAC", "ABCDEFGHIJ", 11, 202, 11))
    p = oracle_lib.params(read_length=12)
    g, r = ctypes.c_void_p(), ctypes.c_void_p()
    assert o.orc_genome_load(_b(fa), None, ctypes.byref(g)) == 0
    try:
        assert o.orc_reads_load(_b(so), ctypes.byref(p), g, ctypes.byref(r)) == 2
    finally:
        o.orc_genome_free(g)


# --------------------------------------------------------- calibration matrix
def matrix_kat_inputs(root: Path):
    """A 500 bp reference (A at 0-199 and 400-499, C at 200-399, N at 150, a
    known SNP at 170) and reads of length 4 whose observed bases (quality
    character not below '@') are placed on purpose; '?' marks a base below
    q_min, which is not observed.  Returns (fasta, soap, dbsnp)."""
    ref = ["A"] * 200 + ["C"] * 200 + ["A"] * 100
    ref[150] = "N"
    fa = root / "m.fa"
    fa.write_text(">chr1\n" + "".join(ref) + "\n")
    snp = root / "m.snp"
    snp.write_text("chr1\t171\t0\t1\t0\t0.5\t0.5\t0.0\t0.0\trs1\n")
    rows = []  # (pos0, seq, qual, strand, hit)
    pos = 0
    for _ in range(12):
        rows.append((pos, "AAAA", "I???", "+", 1))  # q9 cycle 0 A->A
        pos += 1
    for _ in range(10):
        rows.append((pos, "CAAA", "I???", "+", 1))  # q9 cycle 0 A->C
        pos += 1
    for _ in range(5):
        rows.append((pos, "AAAA", "?I??", "+", 1))  # q9 cycle 1 A->A
        pos += 1
    rows.append((pos, "ACAA", "?I??", "+", 1))  # q9 cycle 1 A->C
    pos += 1
    for _ in range(3):
        rows.append((pos, "AATA", "??I?", "+", 1))  # q9 cycle 2 A->T
        pos += 1
    rows.append((pos, "AAAA", "I???", "-", 1))  # reverse strand: coord 0 is cycle 3, A->A
    pos += 1
    # observed bases that must not be counted
    rows.append((pos, "CAAA", "I???", "+", 2))   # two hits
    rows.append((pos, "NAAA", "I???", "+", 1))   # N read base
    rows.append((148, "AACA", "??I?", "+", 1))   # coord 2 lies on the reference N at 150
    rows.append((168, "AACA", "??I?", "+", 1))   # coord 2 lies on the known SNP at 170
    for _ in range(5):
        rows.append((210, "CCCC", "T???", "+", 1))  # q20 cycle 0 C->C
    rows.append((496, "CAAA", "I???", "+", 1))   # pos + len reaches the chromosome end: skipped
    rows.sort(key=lambda t: t[0])
    so = root / "m.soap"
    so.write_text("".join(f"r{i}\t{s}\t{q}\t{h}\ta\t4\t{st}\tchr1\t{p0 + 1}\t0\n" for i, (p0, s, q, st, h) in enumerate(rows)))
    return fa, so, snp


def matrix_kat_expected():
    """Entries of p_matrix[q][cycle][t][o] (A0 C1 T2 G3) for the inputs above,
    worked out from cpp/matrix.cc:129-205 with read_length 4, q_min '@'.

    Counts at q 9: (c0,A,A) 12, (c0,A,C) 10, (c1,A,A) 5, (c1,A,C) 1,
    (c2,A,T) 3, (c3,A,A) 1.  So by type: AA 18, AC 11, AT 3; total 32;
    mismatches 14; per cycle sum over ref A: c0 22, c1 6, c2 3, c3 1.
    10 counts are not "statistically powerful" (the test is > 10).
    Counts at q 20: (c0,C,C) 5; total 5, no mismatches."""
    A, C, T, G = 0, 1, 2, 3
    e = {}
    # q 9, cycle 0: more than 10 counts -> per-cycle ratio over the ref-A sum
    e[(9, 0, A, A)] = 12 / 22
    # (A,C) at cycle 0: exactly 10 counts (not > 10) -> per type: AC 11 over ref-A total 32
    e[(9, 0, A, C)] = 11 / 32
    # (A,T) at cycle 0: 0 counts, AT type 3 -> mismatch rate
    e[(9, 0, A, T)] = 14 / 32
    e[(9, 0, A, G)] = 14 / 32
    # ref C, T, G rows at q 9: no counts of their own types -> match / mismatch rate
    for t in (C, T, G):
        for o in range(4):
            e[(9, 0, t, o)] = 18 / 32 if o == t else 14 / 32
    # cycle 1: (A,A) 5 counts -> per type: AA 18 over 32; (A,C) 1 count -> AC 11 over 32
    e[(9, 1, A, A)] = 18 / 32
    e[(9, 1, A, C)] = 11 / 32
    e[(9, 1, A, T)] = 14 / 32
    # cycle 2: (A,T) 3 counts, type 3 -> mismatch rate
    e[(9, 2, A, T)] = 14 / 32
    e[(9, 2, A, A)] = 18 / 32
    # cycle 3: the reverse-strand base; 1 count -> per type
    e[(9, 3, A, A)] = 18 / 32
    e[(9, 3, A, C)] = 11 / 32
    # q 20: only matches -> rates 1 and 0 -> Phred values 1 - 10^-2 and 10^-2 / 3
    for c in range(4):
        for t in range(4):
            for o in range(4):
                e[(20, c, t, o)] = 1 - math.pow(10, -(20 / 10.0)) if o == t else math.pow(10, -(20 / 10.0)) / 3
    # no data at q 0, 1, 2: the initial 1.0 becomes the Phred value, bounded by 0.25
    for c in range(4):
        for t in range(4):
            for o in range(4):
                e[(0, c, t, o)] = 0.25                      # 1 - 10^0 = 0 -> 0.25; 1/3 -> 0.25
                e[(1, c, t, o)] = 0.25                      # 0.2057 -> 0.25; 0.2642 -> 0.25
                e[(2, c, t, o)] = (1 - math.pow(10, -0.2)) if o == t else math.pow(10, -0.2) / 3
    # cycles at or past read_length and qualities above q_max - q_min (40) keep 1.0
    for t in range(4):
        for o in range(4):
            e[(9, 4, t, o)] = 1.0
            e[(41, 0, t, o)] = 1.0
    return e


def check_matrix_kat(pm: np.ndarray) -> None:
    pm = pm.reshape(256, 256, 4, 4)
    for (q, c, t, o), want in matrix_kat_expected().items():
        got = pm[q, c, t, o]
        assert got == want, f"p_matrix[q={q}][cycle={c}][{'ACTG'[t]}->{'ACTG'[o]}] = {got!r}, want {want!r}"


def test_matrix_gen_fallbacks_by_hand(tmp_path):
    fa, so, snp = matrix_kat_inputs(tmp_path)
    check_matrix_kat(oracle_lib.matrix_gen(fa, so, dbsnp=snp, read_length=4))


# ------------------------------------------------------------------ .cns rows
def _site(**kw):
    s = oracle_lib.OrcSite()
    for k, v in kw.items():
        if isinstance(v, (list, tuple)):
            arr = getattr(s, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(s, k, v)
    return s


def _row(s, chrom="chrX", pos1=1234):
    buf = ctypes.create_string_buffer(4096)
    n = oracle_lib.orc().orc_format_row(chrom.encode(), pos1, ctypes.byref(s), ctypes.byref(s), buf)
    return buf.raw[:n].decode()


def test_cns_n_row_literal():
    # nflag (reference N, depth 0): the literal of cpp/call_genotype.cc:358-361
    s = _site(nflag=1, ori=4, base1=2, base2=3, depth=0, rank_sum=0.5)
    assert _row(s) == "chrX\t1234\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n"
    # base1 >= 4 without nflag: the same literal (cpp/call_genotype.cc:404-406)
    s = _site(nflag=0, ori=1, base1=4, base2=1, depth=3, repeat=3, rank_sum=1.0)
    assert _row(s, pos1=7) == "chrX\t7\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n"


def test_cns_full_row_columns():
    # ref C (code 1) with the known-SNP bit, het CT (type 6 = 'Y'); base1 C,
    # base2 T; average qualities by integer division: 95/3 = 31, 41/2 = 20;
    # rank sum 0.5 -> "0.500000"; copy number 14/4 = 3.5 -> "3.50000"
    s = _site(ori=8 | 1, type1=6, q_cns=37, base1=1, base2=2, q_sum=[0, 95, 41, 0], cnt_uni=[0, 3, 2, 0],
              cnt_all=[1, 4, 2, 0], depth=4, repeat=14, rank_sum=0.5)
    assert _row(s) == "chrX\t1234\tC\tY\t37\tC\t31\t3\t4\tT\t20\t2\t2\t4\t0.500000\t3.50000\t1\n"


def test_cns_base2_not_observed_form():
    # base2 >= 4: "N\t0\t0\t0" in columns 10-13 (cpp/call_genotype.cc:386-401);
    # hom G (type 15 = 'G') on ref G (code 3), rank sum 1 -> "1.00000", copy number 1
    s = _site(ori=3, type1=15, q_cns=99, base1=3, base2=5, q_sum=[0, 0, 0, 200], cnt_uni=[0, 0, 0, 7],
              cnt_all=[0, 0, 0, 8], depth=8, repeat=8, rank_sum=1.0)
    assert _row(s, pos1=1) == "chrX\t1\tG\tG\t99\tG\t28\t7\t8\tN\t0\t0\t0\t8\t1.00000\t1.00000\t0\n"


def test_cns_zero_average_quality_and_zero_depth():
    # q_sum 0 gives average 0 without dividing (count_uni may be 0);
    # depth 0 gives copy number 255 ("255.000"); a tiny rank sum is "%#g"
    s = _site(ori=2, type1=10, q_cns=0, base1=2, base2=0, q_sum=[0, 0, 0, 0], cnt_uni=[0, 0, 0, 0],
              cnt_all=[2, 0, 0, 0], depth=0, repeat=0, rank_sum=1.5e-05)
    assert _row(s, pos1=99) == "chrX\t99\tT\tT\t0\tT\t0\t0\t0\tA\t0\t0\t2\t0\t1.50000e-05\t255.000\t0\n"


# ------------------------------------------------------------- -C window chunk
def _chunk_sites():
    """Six sites of one window (indices 0-5):
    0  nflag (reference N, depth 0)
    1  ref C + known SNP, het CT, base1 C, base2 T, rank sum 0.5
    2  ref T, hom T, base1 T, base2 not observed (5)
    3  ref G, base1 not observed (4): an N row
    4  ref A, call AG ('R'), base1 G, base2 A with q_sum 0 but count_uni 2
       (column 11 is 0 while column 12 is not: the else-if chain of
       cpp/compress.h:1115-1130 records column 12 and drops column 13)
    5  ref C, hom C, base2 G never unique (count_uni 0) but seen twice:
       columns 11 and 12 are 0, column 13 is 2 (an extra column-13 entry)"""
    s = (oracle_lib.OrcSite * 6)()
    s[0] = _site(nflag=1, ori=4)
    s[1] = _site(ori=8 | 1, type1=6, q_cns=25, base1=1, base2=2, q_sum=[0, 90, 40, 0], cnt_uni=[0, 3, 2, 0],
                 cnt_all=[0, 4, 2, 0], depth=7, repeat=14, rank_sum=0.5)
    s[2] = _site(ori=2, type1=10, q_cns=40, base1=2, base2=5, q_sum=[0, 0, 100, 0], cnt_uni=[0, 0, 4, 0],
                 cnt_all=[0, 0, 5, 0], depth=5, repeat=5, rank_sum=1.0)
    s[3] = _site(ori=3, base1=4, base2=4, depth=2, repeat=2, rank_sum=1.0)
    s[4] = _site(ori=0, type1=3, q_cns=25, base1=3, base2=0, q_sum=[0, 0, 0, 90], cnt_uni=[2, 0, 0, 3],
                 cnt_all=[3, 0, 0, 4], depth=0, repeat=0, rank_sum=1.0)
    s[5] = _site(ori=1, type1=5, q_cns=25, base1=1, base2=3, q_sum=[0, 90, 0, 0], cnt_uni=[0, 3, 0, 0],
                 cnt_all=[0, 4, 0, 2], depth=0, repeat=0, rank_sum=1.0)
    return s


def _u64(*v):
    return struct.pack(f"<{len(v)}Q", *v)


def _i32(*v):
    return struct.pack(f"<{len(v)}i", *v)


def _u32(*v):
    return struct.pack(f"<{len(v)}I", *v)


def _f64(*v):
    return struct.pack(f"<{len(v)}d", *v)


def expected_chunk_files() -> dict:
    """The bytes of each file for the window above at 1-based start 262145,
    assembled from cpp/compress.h and cpp/dict_compression.h by hand."""
    f = {}
    # .1_2: "%lu\n" window count, then "%s\t%lu\t%lu\n" (name, posStart, winSize)
    f[".1_2"] = b"1\nchr7\t262145\t6\n"
    # .3_4_6 (Column_3_4_6::compress/write): col3 = "NCTNAC", col4 = "NYTNRC",
    # col6 = "NCTNGC".  col4 differs at 1 ('Y'), 4 ('R'); col6 at 4 ('G'); N in
    # col3 at 0 and 3, replaced by 'A' before the 2-bit encoding (A0 C1 G2 T3,
    # MSB first, last word right-aligned): A C T A A C -> 0b00_01_11_00_00_01 = 449.
    # param {isAllN, dataSize, winSize, codeLen, nN, n4, n6}; dataSize is the
    # buffer: id4 8 + v4 2 + id6 4 + v6 1 + idN 8 + code 4 = 27 bytes
    f[".3_4_6"] = (_u64(0, 27, 6, 1, 2, 2, 1) + _i32(1, 4) + b"YR" + _i32(4) + b"G" + _i32(0, 3) + _u32(449))

    # .5_7_8_9_14_16 (Column_RLEDict::compress_write): N indicator 0, then per
    # column {numRun, nWordValue, nWordLength, nDictValue, nDictLength,
    # dataSize}, value words, length words, value dictionary, length dictionary.
    def rle(nrun, vwords, lwords, vdict, ldict, vfmt="i"):
        vd = struct.pack(f"<{len(vdict)}{vfmt}", *vdict)
        data = _u32(*vwords) + _u32(*lwords) + vd + _u32(*ldict)
        return _u64(nrun, len(vwords), len(lwords), len(vdict), len(ldict), len(data)) + data

    five = _i32(0)
    # col5 q_cns [0,25,40,0,25,25]: runs 0,25,40,0,25 of lengths 1,1,1,1,2;
    # value dict [0,25,40] (2 bits, 16 per word) -> indices 0,1,2,0,1 ->
    # 0b00_01_10_00_01 = 97; length dict [1,2] (1 bit) -> 0,0,0,0,1 -> 1
    five += rle(5, [97], [1], [0, 25, 40], [1, 2])
    # col7 average quality of base1 [0, 90/3=30, 100/4=25, 0, 90/3=30, 90/3=30]:
    # dict [0,25,30] -> 0,2,1,0,2 -> 0b00_10_01_00_10 = 146
    five += rle(5, [146], [1], [0, 25, 30], [1, 2])
    # col8 count_uni of base1 [0,3,4,0,3,3]: dict [0,3,4] -> 0,1,2,0,1 -> 97
    five += rle(5, [97], [1], [0, 3, 4], [1, 2])
    # col9 count_all of base1 [0,4,5,0,4,4] -> 97
    five += rle(5, [97], [1], [0, 4, 5], [1, 2])
    # col14 depth [0,7,5,0,0,0]: runs 0,7,5,0 of lengths 1,1,1,3; value dict
    # [0,5,7] -> 0,2,1,0 -> 0b00_10_01_00 = 36; length dict [1,3] (1 bit, 32
    # per word) -> 0,0,0,1 -> 1
    five += rle(4, [36], [1], [0, 5, 7], [1, 3])
    # col16 copy number [255, 14/7=2, 5/5=1, 255, 255 (depth 0), 255]: runs
    # 255,2,1,255 of lengths 1,1,1,3; dict [1,2,255] -> 2,1,0,2 -> 146; f64 dictionary
    five += rle(4, [146], [1], [1.0, 2.0, 255.0], [1, 3], vfmt="d")
    f[".5_7_8_9_14_16"] = five
    # .10 (Column_10::compress_write): col10 = base2 where both bases are
    # observed, else 'N': N T N N A G; N at 0, 2, 3 become 'A': A T A A A G ->
    # 0b00_11_00_00_00_10 = 770; param {numN, numWord, numElement}
    f[".10"] = _i32(0) + _u64(3, 1, 6) + _u32(770) + _u32(0, 2, 3)
    # .11_12_13 (Column_11_12_13::compress_write): site 1 has column 11 = 20
    # (common: 20, 2, 2); site 4 has column 11 = 0 and column 12 = 2 (extra
    # column 12; its column 13 = 3 is dropped by the else-if chain); site 5
    # has columns 11 and 12 = 0 and column 13 = 2 (extra column 13).  param
    # {winSize, numCommon, numExtra12, numExtra13}, extra12 ids, values,
    # extra13 ids, values, common ids, then per common column {nWord, nDict},
    # words, dictionary (one value: 1 bit, one word holding index 0)
    f[".11_12_13"] = (_i32(0) + _u64(6, 1, 1, 1) + _i32(4) + _i32(2) + _i32(5) + _i32(2) + _i32(1) +
                      _u64(1, 1) + _u32(0) + _i32(20) +
                      _u64(1, 1) + _u32(0) + _i32(2) +
                      _u64(1, 1) + _u32(0) + _i32(2))
    # .15_17 (Column_15_17::compress_write): rank sums != 1 of called sites
    # (site 1: 0.5), then the known-SNP flags of called sites (site 1)
    f[".15_17"] = _i32(0) + _i32(1) + _i32(1) + _f64(0.5) + _i32(1) + _i32(1)
    return f


def test_c_chunk_bytes_by_hand(tmp_path):
    pre = tmp_path / "kat"
    sites = _chunk_sites()
    assert oracle_lib.orc().orc_cwrite_sites(_b(pre), b"chr7", 262144, 6, sites) == 0
    for sfx, want in expected_chunk_files().items():
        got = Path(f"{pre}{sfx}").read_bytes()
        assert got == want, f"{sfx}: got {got.hex()} want {want.hex()}"


def test_c_chunk_all_n_window_by_hand(tmp_path):
    # a window of nflag sites only: .3_4_6 param {1, 0, winSize, 0, 0, 0, 0},
    # every other file the N indicator 1 alone (cpp/compress.h:335-346, :585-590, :857-862, ...)
    pre = tmp_path / "alln"
    s = (oracle_lib.OrcSite * 3)()
    for i in range(3):
        s[i] = _site(nflag=1, ori=4)
    assert oracle_lib.orc().orc_cwrite_sites(_b(pre), b"c", 0, 3, s) == 0
    assert Path(f"{pre}.1_2").read_bytes() == b"1\nc\t1\t3\n"
    assert Path(f"{pre}.3_4_6").read_bytes() == _u64(1, 0, 3, 0, 0, 0, 0)
    for sfx in (".5_7_8_9_14_16", ".10", ".11_12_13", ".15_17"):
        assert Path(f"{pre}{sfx}").read_bytes() == _i32(1), sfx


# ------------------------------------------------------------- -T region masks
def _mask(len_, ranges):
    o = oracle_lib.orc()
    nw = int(o.orc_region_words(len_))
    site = np.zeros(nw, dtype=np.uint64)
    win = np.zeros(int(o.orc_region_words(len_ // 262144 + 1)), dtype=np.uint64)
    for a, b in ranges:
        assert o.orc_region_set(site.ctypes.data, win.ctypes.data, len_, a, b) == 0
    return site, win


def test_region_site_mask_words_by_hand():
    """Chr_info::set_region (cpp/chromosome.cc:66-107): position p is bit
    63 - p % 64 of word p / 64, so a range fills words from their top bit."""
    site, _ = _mask(200, [(3, 5)])  # one word: bits 60, 59, 58
    assert [int(x) for x in site] == [0x1C00000000000000, 0, 0, 0]
    site, _ = _mask(200, [(60, 66)])  # across a word boundary: 60-63, then 64-66
    assert [int(x) for x in site] == [0xF, 0xE000000000000000, 0, 0]
    site, _ = _mask(200, [(64, 127)])  # a whole word (the reference's shift by 64 sets nothing here on x86)
    assert [int(x) for x in site] == [0, 0xFFFFFFFFFFFFFFFF, 0, 0]
    site, _ = _mask(200, [(-40, 1)])  # clipped at 0
    assert int(site[0]) == 0xC000000000000000
    site, _ = _mask(200, [(190, 900)])  # clipped at the length (the end bit sits at position len)
    assert int(site[2]) == 0x3 and int(site[3]) == 0xFF80000000000000  # 190-191, then 192-200
    site, _ = _mask(200, [(3, 5), (4, 9)])  # ranges accumulate
    assert int(site[0]) == 0x1FC0000000000000


def test_region_window_mask_by_hand():
    """Windows start / 262144 .. end / 262144 of the clipped range."""
    W = 262144
    _, win = _mask(3 * W + 10, [(W - 5, W + 5)])
    assert int(win[0]) == 0xC000000000000000  # windows 0 and 1: bits 63 and 62
    o = oracle_lib.orc()
    assert [o.orc_region_in_win(win.ctypes.data, p) for p in (0, W - 1, W, 2 * W, 3 * W)] == [1, 1, 1, 0, 0]
    _, win = _mask(3 * W + 10, [(3 * W, 3 * W + 9)])  # the last, partial window only
    assert int(win[0]) == 0x1000000000000000


def test_region_start_after_end_is_an_error():
    o = oracle_lib.orc()
    site = np.zeros(4, dtype=np.uint64)
    assert o.orc_region_set(site.ctypes.data, None, 200, 50, 40) == 2
