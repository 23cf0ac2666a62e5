"""C ABI checks that need no GPU: every entry point declared in
include/gsnp_abi.h is exported by libgsnp.so, the host-side entry points
(formatter, seeded generator, calibration matrix, decompressor) behave, and
malformed inputs come back as the documented error codes."""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gsnp_abi.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsnp_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported(built):
    so = ctypes.CDLL(str(built.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    # and the Python binding types every one of them
    assert set(names) <= set(built.SIGNATURES), sorted(set(names) - set(built.SIGNATURES))


def test_format_g_matches_printf(built):
    """The device formatter's host build against the C library's printf
    (Python's own %-formatting differs from glibc at the carry into a new
    decade, e.g. 999999.5 prints as 1.e+06 under glibc)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.snprintf.restype = ctypes.c_int
    rng = np.random.default_rng(5)
    vals = list(rng.uniform(0, 1, 300)) + list(rng.uniform(0, 300, 300)) + list(10 ** rng.uniform(-8, 8, 300))
    vals += [0.0, 1.0, 255.0, 1e-5, 999999.5, 999999.6, 99999.95, 0.5, 1.0 / 3, 2.0 / 3, 99.99995]
    buf = ctypes.create_string_buffer(64)
    for v in vals:
        libc.snprintf(buf, 64, b"%#g", ctypes.c_double(float(v)))
        assert built.format_g(float(v)) == buf.value.decode(), v


def test_generator_is_deterministic(built, tmp_path):
    a = built.synth(1, 99, 0.05, tmp_path / "a")
    b = built.synth(1, 99, 0.05, tmp_path / "b")
    assert a == b
    for ext in (".fa", ".soap"):
        assert (tmp_path / ("a" + ext)).read_bytes() == (tmp_path / ("b" + ext)).read_bytes()
    c = built.synth(1, 100, 0.05, tmp_path / "c")
    assert (tmp_path / "a.soap").read_bytes() != (tmp_path / "c.soap").read_bytes() or c != a


def _soap_line(name, pos1, seq, qual="I", strand="+", hits=1, chrom="chr1"):
    q = qual * len(seq) if len(qual) == 1 else qual
    return f"{name}\t{seq}\t{q}\t{hits}\ta\t{len(seq)}\t{strand}\t{chrom}\t{pos1}\t0\n"


def _write_inputs(tmp_path, lines, ref="ACGT" * 100):
    fa = tmp_path / "r.fa"
    fa.write_text(">chr1\n" + ref + "\n")
    so = tmp_path / "r.soap"
    so.write_text("".join(lines))
    return fa, so


def test_matrix_gen_rejects_unsorted_alignments(built, tmp_path):
    fa, so = _write_inputs(tmp_path, [_soap_line("r1", 50, "ACGTACGTAC"), _soap_line("r2", 10, "GTACGTACGT")])
    pm = np.zeros(256 * 256 * 16)
    p = built.params(read_length=10)
    rc = built.lib().gsnp_matrix_gen(os.fsencode(fa), os.fsencode(so), None, ctypes.byref(p), pm.ctypes.data)
    assert rc == 3, built.lib().gsnp_last_error()


def test_matrix_gen_rejects_unknown_chromosome(built, tmp_path):
    fa, so = _write_inputs(tmp_path, [_soap_line("r1", 5, "ACGTACGTAC", chrom="chrX")])
    pm = np.zeros(256 * 256 * 16)
    p = built.params(read_length=10)
    rc = built.lib().gsnp_matrix_gen(os.fsencode(fa), os.fsencode(so), None, ctypes.byref(p), pm.ctypes.data)
    assert rc == 4, built.lib().gsnp_last_error()


@pytest.mark.parametrize("case", ["unsorted", "unknown", "noncontiguous", "unknown_then_unsorted"])
def test_threaded_reader_reports_the_first_error(built, tmp_path, monkeypatch, case):
    """The SOAP reader parses slices of the file on threads and merges them in
    file order: with one slice or seven, the same first error is reported."""
    ref = "ACGT" * 2500
    lines = [_soap_line(f"r{i}", 1 + i, ref[i:i + 10]) for i in range(3000)]
    if case == "unsorted":
        lines[2500], lines[2501] = lines[2501], lines[2500]
    elif case == "unknown":
        lines[1700] = _soap_line("x", 1701, ref[1700:1710], chrom="chrZ")
    elif case == "noncontiguous":
        lines = [_soap_line("c2", 5, ref[4:14], chrom="chr2")] + lines + [_soap_line("c2b", 9, ref[8:18], chrom="chr2")]
    else:
        lines[400] = _soap_line("x", 401, ref[400:410], chrom="chrZ")
        lines[2800], lines[2801] = lines[2801], lines[2800]
    fa = tmp_path / "r.fa"
    fa.write_text(">chr1\n" + ref + "\n>chr2\n" + ref + "\n")
    so = tmp_path / "r.soap"
    so.write_text("".join(lines))
    got = []
    for slices in ("1", "7"):
        monkeypatch.setenv("GSNP_PARSE_THREADS", slices)
        pm = np.zeros(256 * 256 * 16)
        p = built.params(read_length=10)
        rc = built.lib().gsnp_matrix_gen(os.fsencode(fa), os.fsencode(so), None, ctypes.byref(p), pm.ctypes.data)
        got.append((rc, built.lib().gsnp_last_error()))
    assert got[0] == got[1]
    assert got[0][0] in (3, 4), got[0]


@pytest.mark.gpu
def test_matrix_gen_counts_a_clean_read(built, tmp_path):
    ref = "ACGT" * 100
    fa, so = _write_inputs(tmp_path, [_soap_line(f"r{i}", 1 + 4 * i, ref[4 * i:4 * i + 10]) for i in range(60)], ref)
    pm = built.matrix_gen(fa, so, read_length=10).reshape(256, 256, 4, 4)
    q = ord("I") - 64  # quality 9
    # no mismatches were seen: every cycle falls back to the Phred value
    e = 10 ** (-q / 10.0)
    assert pm[q, 0, 0, 0] == pytest.approx(1 - e, rel=0, abs=0)
    assert pm[q, 0, 0, 1] == pytest.approx(e / 3, rel=0, abs=0)


def test_decompressor_reports_missing_files(built, tmp_path):
    rc = built.lib().gsnp_decompress(os.fsencode(tmp_path / "nothing"), os.fsencode(tmp_path / "out.cns"))
    assert rc == 1


@pytest.mark.parametrize("name", ["c1", "stress"])
def test_cns_reader_gives_the_oracle_rows(built, datasets, tmp_path, name):
    """-C files written by the oracle, read in memory window by window
    (gsnp_cns_open / gsnp_cns_next), give the oracle's .cns rows column for
    column; stress has three chromosomes and known SNPs."""
    import oracle_lib

    ds = datasets(name)
    prefix = tmp_path / name
    oracle_lib.run(ds.fasta, ds.soap, prefix, dbsnp=ds.dbsnp, compress=True, **ds.flags)
    rows = []
    for chrom, start, c in built.cns_windows(prefix):
        for i in range(len(c["ref"])):
            ref, cns, b1 = chr(c["ref"][i]), chr(c["cns"][i]), chr(c["base1"][i])
            if ref == cns == b1 == "N":
                rows.append(f"{chrom}\t{start + i}\tN\tN\t0\tN\t0\t0\t0\tN\t0\t0\t0\t0\t1.000\t255.000\t0\n")
                continue
            rows.append("\t".join([chrom, str(start + i), ref, cns, str(c["q_cns"][i]), b1, str(c["avg_q1"][i]),
                                   str(c["count_uni1"][i]), str(c["count_all1"][i]), chr(c["base2"][i]),
                                   str(c["avg_q2"][i]), str(c["count_uni2"][i]), str(c["count_all2"][i]),
                                   str(c["depth"][i]), "%#g" % c["rank_sum"][i], "%#g" % c["copy_num"][i],
                                   str(c["dbsnp"][i])]) + "\n")
    want = ds.oracle_text().read_text().splitlines(keepends=True)
    assert len(rows) == len(want)
    bad = [i for i in range(len(rows)) if rows[i] != want[i]]
    assert not bad, f"{len(bad)} rows differ, first {bad[0]}: {rows[bad[0]]!r} vs {want[bad[0]]!r}"


def test_cns_reader_reports_missing_files(built, tmp_path):
    r = ctypes.c_void_p()
    assert built.lib().gsnp_cns_open(os.fsencode(tmp_path / "nothing"), ctypes.byref(r)) == 1
