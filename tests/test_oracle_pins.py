"""Pins of the CPU oracle (CPU only).

The reference ships no golden vectors for this path, so each oracle function
is pinned by a check a plausible mistake would fail: hand-computed known
answers, closed forms, brute force on small inputs, an independent second
statement (dense vs sorted likelihood), and the reference's own dictionary
codec compiled from cpp/dict_compression.h (oracle/_ref).
"""
from __future__ import annotations

import ctypes
import itertools
import math

import numpy as np
import pytest

import oracle_lib

o = None


@pytest.fixture(scope="module", autouse=True)
def lib():
    global o
    o = oracle_lib.orc()
    return o


def enc(coord, strand, base, q, occ=1):
    return (base << 23) | ((64 - q) << 17) | (coord << 9) | (strand << 8) | occ


# ---------------------------------------------------------------- record codec
def test_record_codec_layout():
    # kernel/likelihood_kernel.cu:24-42 closed form
    assert o.orc_encode_word(5, 1, 2, 30, 1) == (2 << 23) | (34 << 17) | (5 << 9) | (1 << 8) | 1
    assert o.orc_encode_word(0, 0, 0, 63, 1) == (1 << 17) | 1
    assert o.orc_encode_word(255, 1, 3, 1, 1) == (3 << 23) | (63 << 17) | (255 << 9) | (1 << 8) | 1
    assert o.orc_encode_word(255, 1, 3, 1, 1) == 0x01FFFF01


def test_sort_order_is_base_then_quality_descending():
    # ascending u32 = base up, q down, cycle up, strand up
    recs = [enc(3, 0, 1, 20), enc(9, 1, 0, 10), enc(2, 0, 0, 35), enc(2, 1, 0, 35), enc(1, 0, 0, 35)]
    arr = (ctypes.c_uint32 * len(recs))(*recs)
    o.orc_sort_records(arr, len(recs))
    got = [(r >> 23 & 3, 64 - (r >> 17 & 63), r >> 9 & 255, r >> 8 & 1) for r in arr]
    assert got == [(0, 35, 1, 0), (0, 35, 2, 0), (0, 35, 2, 1), (0, 10, 9, 1), (1, 20, 3, 0)]


# ----------------------------------------------------------------------- q_adj
def test_qadj_closed_forms():
    l05, l09 = math.log10(0.5), math.log10(0.9)
    # first occurrence of the first key is the raw quality
    for q in range(1, 64):
        assert o.orc_qadj(q, 1, 0, l05, l09) == q
    # hand values at the default -p 0.5 -g 0.9
    assert o.orc_qadj(30, 2, 0, l05, l09) == 15   # 30 * 0.5
    assert o.orc_qadj(30, 1, 1, l05, l09) == 27   # 30 * 0.9
    assert o.orc_qadj(30, 1, 2, l05, l09) == 24   # 24.3
    assert o.orc_qadj(20, 3, 2, l05, l09) == 4    # 20 * .25 * .81 = 4.05
    assert o.orc_qadj(2, 3, 9, l05, l09) == 1     # floor at 1
    # -p 1 -g 1: no dependency, q_adj = q
    for q in range(1, 64):
        for t, g in [(1, 0), (5, 0), (1, 7), (9, 30)]:
            assert o.orc_qadj(q, t, g, 0.0, 0.0) == q


# ----------------------------------------------------------------- score table
def _score_table(pm):
    s = np.zeros(256 * 256 * 4 * 10)
    o.orc_score_table(pm.ctypes.data, s.ctypes.data)
    return s


def test_score_table_genotype_order():
    # identity calibration: p[q][c][a][o] = [a == o]; then
    # S(g = a1 a2, o) = log10(0.5[a1==o] + 0.5[a2==o]): 0 hom match, log10(.5) het, -inf otherwise
    pm = np.zeros(256 * 256 * 16)
    for a in range(4):
        pm[np.arange(256 * 256) * 16 + a * 4 + a] = 1.0
    s = _score_table(pm)
    order = [(a1, a2) for a1 in range(4) for a2 in range(a1, 4)]
    assert order == [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
    for q, c in [(0, 0), (30, 17), (255, 255)]:
        for ob in range(4):
            row = s[((q << 10) | (c << 2) | ob) * 10:][:10]
            for k, (a1, a2) in enumerate(order):
                hits = (a1 == ob) + (a2 == ob)
                want = 0.0 if hits == 2 else (math.log10(0.5) if hits == 1 else -math.inf)
                assert row[k] == want, (q, c, ob, a1, a2)


# ----------------------------------------------------------------- likelihood
def _lk(recs, fn="sorted", L=100, pcr=math.log10(0.5), glob=math.log10(0.9), score=None, qmm=40):
    arr = np.ascontiguousarray(np.array(recs, dtype=np.uint32))
    out = np.zeros(10)
    if fn == "sorted":
        rc = o.orc_likelihood_sorted(arr.ctypes.data, len(arr), L, pcr, glob, score.ctypes.data, out.ctypes.data)
    else:
        rc = o.orc_likelihood_dense(arr.ctypes.data, len(arr), L, qmm, pcr, glob, score.ctypes.data, out.ctypes.data)
    assert rc == 0
    return out


@pytest.fixture(scope="module")
def score():
    rng = np.random.default_rng(3)
    pm = rng.uniform(0.001, 1.0, 256 * 256 * 16)
    return _score_table(pm)


def test_single_record_is_the_score_row(score):
    for w in [enc(5, 0, 2, 30), enc(99, 1, 0, 2), enc(0, 0, 3, 63)]:
        q, c, b = 64 - (w >> 17 & 63), w >> 9 & 255, w >> 23 & 3
        row = score[((q << 10) | (c << 2) | b) * 10:][:10]
        assert np.array_equal(_lk([w], score=score), row)


def test_dependency_counts_by_hand(score):
    # same base; keys (strand 0, cycle 5) twice and (strand 1, cycle 5) once
    l05, l09 = math.log10(0.5), math.log10(0.9)
    recs = sorted([enc(5, 0, 1, 30), enc(5, 0, 1, 20), enc(5, 1, 1, 25)])
    # sorted: q30 (k0), q25 (k1), q20 (k0 again)
    t_g = [(30, 1, 0), (25, 1, 1), (20, 2, 1)]
    want = np.zeros(10)
    for (q, t, g), w in zip(t_g, recs):
        qa = o.orc_qadj(q, t, g, l05, l09)
        want += score[((qa << 10) | ((w >> 9 & 255) << 2) | 1) * 10:][:10]
    assert [o.orc_qadj(*x, l05, l09) for x in t_g] == [30, 23, 9]
    assert np.array_equal(_lk(recs, score=score), want)


def test_sorted_and_dense_statements_agree(score):
    rng = np.random.default_rng(7)
    for trial in range(300):
        n = int(rng.integers(1, 256))
        base = rng.integers(0, 4, n)
        q = rng.integers(1, 41, n)
        cyc = rng.integers(0, 6 if trial % 2 else 100, n)
        st = rng.integers(0, 2, n)
        recs = sorted(int(x) for x in (base << 23) | ((64 - q) << 17) | (cyc << 9) | (st << 8) | 1)
        a = _lk(recs, "sorted", score=score)
        b = _lk(recs, "dense", score=score)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), trial


# ---------------------------------------------------------------------- priors
def test_prior_table_by_hand():
    p = oracle_lib.params()
    pp = np.zeros(128)
    o.orc_prior_gen(ctypes.byref(p), pp.ctypes.data)
    het, alt = 0.001, 0.0005
    # ref A (code 0): genotypes a1<=a2 over A0 C1 T2 G3
    want_A = {(0, 0): 1, (0, 1): het, (0, 2): het, (0, 3): het, (1, 1): alt, (1, 2): het * alt, (1, 3): het * alt,
              (2, 2): alt, (2, 3): het * alt, (3, 3): alt}
    for (a1, a2), v in want_A.items():
        assert pp[0 << 4 | a1 << 2 | a2] == v
    # N reference rows
    for r in range(4, 8):
        for a1 in range(4):
            for a2 in range(a1, 4):
                assert pp[r << 4 | a1 << 2 | a2] == (1 if a1 == a2 else 2 * het) * 0.25 * 0.25
    # transition-dominant: genotypes holding the transition partner (A<->G, code ^ 3) are x4
    p.transition_dominant = 1
    o.orc_prior_gen(ctypes.byref(p), pp.ctypes.data)
    for (a1, a2), v in want_A.items():
        tr = (a1 ^ 0) == 3 or (a2 ^ 0) == 3
        assert pp[a1 << 2 | a2] == v * (4 if tr else 1)


def test_known_snp_prior_branches():
    p = oracle_lib.params()
    row = np.full(16, -1.0)
    s = oracle_lib.OrcSnp(0, 1, 0, (ctypes.c_double * 4)(0.7, 0.0, 0.3, 0.0))  # A/T, validated
    o.orc_snp_prior(row.ctypes.data, 0, ctypes.byref(s), ctypes.byref(p))
    assert row[0 << 2 | 0] == 1 and row[0 << 2 | 2] == p.het_val_r and row[2 << 2 | 2] == p.althom_val_r
    assert row[1 << 2 | 1] == -1.0  # untouched
    row[:] = -1
    s = oracle_lib.OrcSnp(1, 0, 0, (ctypes.c_double * 4)(0.7, 0.0, 0.3, 0.0))  # HapMap frequencies
    o.orc_snp_prior(row.ctypes.data, 0, ctypes.byref(s), ctypes.byref(p))
    assert row[0] == 0.7 * 0.7 and row[2] == 2 * p.het_val_r * 0.7 * 0.3 and row[10] == 0.3 * 0.3
    row[:] = -1
    s = oracle_lib.OrcSnp(0, 0, 1, (ctypes.c_double * 4)(0.7, 0.0, 0.3, 0.0))  # indel: no change
    o.orc_snp_prior(row.ctypes.data, 0, ctypes.byref(s), ctypes.byref(p))
    assert (row == -1).all()


# ------------------------------------------------------------------ rank table
@pytest.fixture(scope="module")
def prank():
    pr = np.zeros(64 * 64 * 2048)
    o.orc_rank_table_gen(pr.ctypes.data)
    return pr


def test_rank_table_by_subset_enumeration(prank):
    for N in range(1, 15):
        for n1 in range(0, N + 1):
            counts = {}
            for sub in itertools.combinations(range(1, N + 1), n1):
                counts[sum(sub)] = counts.get(sum(sub), 0) + 1
            total = math.comb(N, n1)
            lo, hi = n1 * (n1 + 1) // 2, (2 * N - n1 + 1) * n1 // 2
            left = 0.0
            for T in range(lo, hi + 1):
                right = 1.0 - left
                left += counts.get(T, 0) / total
                want = min(left, right)
                assert prank[N << 17 | n1 << 11 | T] == pytest.approx(want, rel=1e-12, abs=1e-15), (N, n1, T)


def test_rank_table_symmetric_and_bounded(prank):
    for N in (20, 40, 63):
        for n1 in (1, 5, N // 2):
            lo, hi = n1 * (n1 + 1) // 2, (2 * N - n1 + 1) * n1 // 2
            vals = [prank[N << 17 | n1 << 11 | T] for T in range(lo, hi + 1)]
            # min(P(X<=T), P(X>=T)) exceeds 0.5 only at the centre; the far right tail
            # (1 - left sum) cancels to 0 for large N, as cpp/rank_sum.cc:30-35 computes it
            assert all(0.0 <= v <= 1.0 for v in vals)
            assert vals == pytest.approx(vals[::-1], rel=0, abs=1e-13)  # right tail is 1 - (left sum)


def test_table_test_interpolation_index(prank):
    # cpp/rank_sum.cc:51-58 interpolates toward index (n1+n2)<<16|n1<<11|(int)(T+1)
    pr = np.arange(64 * 64 * 2048, dtype=np.float64) * 1e-9
    n1, n2, T1 = 5, 9, 31.5
    i1 = (n1 + n2) << 17 | n1 << 11 | 31
    i2 = (n1 + n2) << 16 | n1 << 11 | 32
    want = pr[i1] + 0.5 * (pr[i2] - pr[i1])
    assert o.orc_table_test(pr.ctypes.data, n1, n2, T1, 60.0) == want
    # n1 > n2 uses n2 and T2
    i1 = (n1 + n2) << 17 | n2 << 11 | 40
    assert o.orc_table_test(pr.ctypes.data, 12, 5, 9.0, 40.0) != 0
    assert o.orc_table_test(pr.ctypes.data, n2, n1, 60.0, 40.0) == pr[(n1 + n2) << 17 | n1 << 11 | 40]


def test_normal_value_against_erfc():
    for z in np.linspace(-6, 6, 2001):
        tail = 0.5 * math.erfc(abs(z) / math.sqrt(2))
        assert abs(o.orc_normal_value(z) - tail) < 7.5e-8 + 1e-15
    assert o.orc_normal_value(6.01) == 0.0 and o.orc_normal_value(-7.0) == 0.0


def test_rank_test_brute_force(prank):
    # mid-rank sums over the two alleles' qualities vs a direct statement
    rng = np.random.default_rng(11)
    for _ in range(200):
        n = int(rng.integers(2, 60))
        base = rng.choice([0, 2], n)
        base[0], base[1] = 0, 2
        q = rng.integers(1, 41, n)
        recs = sorted(int(x) for x in (base << 23) | ((64 - q) << 17) | (1 << 9) | 1)
        cu = (ctypes.c_int * 4)(int((base == 0).sum()), 0, int((base == 2).sum()), 0)
        best = 0 << 2 | 2  # het A/T: n1 = count of T (best&3 = 2), n2 = count of A
        got = o.orc_rank_test((ctypes.c_uint32 * n)(*recs), n, cu, 40, best, prank.ctypes.data)
        # direct: ranks by quality ascending, ties get mid-ranks
        order = np.sort(q)
        T = {0: 0.0, 2: 0.0}
        for bb, qq in zip(base, q):
            lo = int(np.searchsorted(order, qq, "left")) + 1
            hi = int(np.searchsorted(order, qq, "right"))
            T[int(bb)] += (lo + hi) / 2.0
        n1, n2 = cu[2], cu[0]
        if n1 + n2 < 64:
            want = o.orc_table_test(prank.ctypes.data, n1, n2, T[2], T[0])
        else:
            want = o.orc_normal_test(n1, n2, T[2], T[0])
        assert got == want


# --------------------------------------------------------- posterior branches
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


def _post(s, recs=(), nxt=(0, 0, 0, 0), **pkw):
    p = oracle_lib.params(**pkw)
    pp = np.zeros(128)
    o.orc_prior_gen(ctypes.byref(p), pp.ctypes.data)
    pr = np.zeros(64 * 64 * 2048)
    o.orc_rank_table_gen(pr.ctypes.data)
    arr = (ctypes.c_uint32 * max(1, len(recs)))(*recs)
    nq = (ctypes.c_int32 * 4)(*nxt)
    o.orc_posterior_site(ctypes.byref(s), arr, len(recs), nq, ctypes.byref(p), pp.ctypes.data, None, pr.ctypes.data)
    return s


def test_posterior_uncovered_site():
    # depth 0 on reference C: base1 = ref, base2 = T (later bases win >= ties), q_cns capped at qual1-qual2 = 1
    s = _post(_site(ori=1, lk=[0.0] * 10))
    assert (s.base1, s.base2, s.type1, s.q_cns, s.rank_sum) == (1, 2, 5, 1, 1.0)


def test_posterior_hom_alt_and_caps():
    # 10 unique T (code 2) observations on reference A; likelihood favours TT strongly
    lk = [-30.0] * 10
    lk[7] = -1.0  # TT
    s = _post(_site(ori=0, depth=10, dep_uni=10, cnt_uni=[0, 0, 10, 0], q_sum=[0, 0, 300, 0],
                    cnt_all=[0, 0, 10, 0], lk=lk))
    assert s.base1 == 2 and s.type1 == 10  # TT
    # q_cns = 10*(tp1 - tp2) capped at qual1 - qual2 = 300 - 0, then at 99
    assert s.q_cns == 99


def test_posterior_het_requires_both_alleles_observed():
    lk = [-30.0] * 10
    lk[1] = -2.0  # AC
    s = _post(_site(ori=0, depth=8, dep_uni=8, cnt_uni=[8, 0, 0, 0], q_sum=[240, 0, 0, 0], cnt_all=[8, 0, 0, 0],
                    lk=lk))
    assert s.type1 == 1 and s.q_cns == 0  # HET with one allele observed


def test_posterior_monoploid_calls_only_homs():
    lk = [-1.0] * 10
    lk[1] = 0.0
    s = _post(_site(ori=0, depth=8, dep_uni=8, cnt_uni=[4, 4, 0, 0], q_sum=[100, 100, 0, 0], cnt_all=[4, 4, 0, 0],
                    lk=lk), is_monoploid=1)
    assert (s.type1 & 3) == (s.type1 >> 2 & 3)


def test_posterior_n_reference_het_next_site_is_unobservable():
    # N reference ('N' binarizes to code 7) with only C observed: base2 becomes the
    # N code and the HET cap reads q_sum through the flat index (the next site's
    # G sum, kernel/posterior_kernel.cu:848-851).  base2 = N needs qual2 == 0, so
    # the cap qual2 - qual3 is <= 0 and q_cns ends at 0 whatever the next site holds.
    lk = [-30.0] * 10
    lk[6] = -0.5  # CG
    base = dict(ori=7, depth=6, dep_uni=6, cnt_uni=[0, 6, 0, 0], q_sum=[0, 180, 0, 0], cnt_all=[0, 6, 0, 0], lk=lk)
    for nxt in [(0, 0, 0, 0), (0, 0, 0, 5), (9, 9, 9, 9)]:
        s = _post(_site(**base), nxt=nxt)
        assert (s.base1, s.base2, s.type1, s.q_cns) == (1, 7, 7, 0)


# --------------------------------------------------------------------- text
@pytest.mark.parametrize("v,s", [(1.0, "1.00000"), (255.0, "255.000"), (1.5e-5, "1.50000e-05"), (0.0, "0.00000"),
                                 (0.000123456789, "0.000123457"), (0.474858, "0.474858"), (1.0833333333, "1.08333"),
                                 (999999.5, "1.e+06"), (9999995.0, "1.00000e+07")])
def test_format_g_kats(v, s):
    buf = ctypes.create_string_buffer(64)
    n = o.orc_format_g(v, buf)
    assert buf.raw[:n].decode() == s


# ------------------------------------------------------------------ -C codec
def test_num_bits_matches_reference():
    ref = oracle_lib.dictref()
    for n in list(range(1, 4200)) + [2 ** k + d for k in range(12, 31) for d in (-1, 0, 1)]:
        b = o.orc_num_bits(n)
        assert b == max(1, math.ceil(math.log(n) / math.log(2.0)))
        if ref is not None:
            assert b == ref.ref_num_bits(n)


def test_dictionary_encoding_matches_reference_codec():
    ref = oracle_lib.dictref()
    if ref is None:
        pytest.skip("reference tree not present (oracle/_ref not built)")
    rng = np.random.default_rng(21)
    cases = [np.array([7], np.int32), np.array([3, 1, 3, 7, 1], np.int32)]
    cases += [rng.integers(-k, k + 1, int(rng.integers(1, 3000))).astype(np.int32) for k in (1, 2, 3, 7, 8, 100, 5000)]
    for v in cases:
        n = len(v)
        w1, d1, nd1 = np.zeros(n + 1, np.uint32), np.zeros(n + 1, np.int32), ctypes.c_uint64()
        nw1 = o.orc_dict_encode_i32(v.ctypes.data, n, w1.ctypes.data, d1.ctypes.data, ctypes.byref(nd1))
        w2, d2, nd2 = np.zeros(n + 1, np.uint32), np.zeros(n + 1, np.int32), ctypes.c_ulong()
        nw2 = ref.ref_dict_encode_i32(v.ctypes.data, n, w2.ctypes.data, d2.ctypes.data, ctypes.byref(nd2))
        assert (nw1, nd1.value) == (nw2, nd2.value)
        assert np.array_equal(w1[:nw1], w2[:nw2]) and np.array_equal(d1[:nd1.value], d2[:nd2.value])
        back = np.zeros(n, np.int32)
        ref.ref_dict_decode_i32(back.ctypes.data, n, w1.ctypes.data, nw1, d1.ctypes.data, nd1.value)
        assert np.array_equal(back, v)
    # doubles (column 16)
    v = rng.choice([255.0, 1.0, 1.5, 2.0, 1.0 / 3.0], 777)
    w1, d1, nd1 = np.zeros(778, np.uint32), np.zeros(778), ctypes.c_uint64()
    nw1 = o.orc_dict_encode_f64(v.ctypes.data, 777, w1.ctypes.data, d1.ctypes.data, ctypes.byref(nd1))
    w2, d2, nd2 = np.zeros(778, np.uint32), np.zeros(778), ctypes.c_ulong()
    nw2 = ref.ref_dict_encode_f64(v.ctypes.data, 777, w2.ctypes.data, d2.ctypes.data, ctypes.byref(nd2))
    assert nw1 == nw2 and nd1.value == nd2.value and np.array_equal(w1[:nw1], w2[:nw2])


def test_two_bit_codec():
    code = (ctypes.c_uint32 * 2)()
    assert o.orc_encode_2bit(b"ACGT", 4, code) == 1 and code[0] == 0b00011011
    seq = b"ACGTTGCAACGTTGCAG"  # 17 bases: second word holds one right-aligned base
    assert o.orc_encode_2bit(seq, 17, code) == 2
    assert code[0] == 0x1BE41BE4 and code[1] == 2
    out = ctypes.create_string_buffer(17)
    o.orc_decode_2bit(out, 17, code, 2)
    assert out.raw == seq


# ------------------------------------------------- window-parallel oracle run
def test_threaded_run_equals_serial_run(tmp_path):
    """The CPU baseline's window-parallel run (-J) writes the serial run's text:
    each window is called from every read that covers it, independently."""
    import subprocess

    import gsnp_amd

    pre = tmp_path / "c2"
    gsnp_amd.synth(2, 12, 0.06, pre)  # 600 kb: three windows
    cli = str(oracle_lib.ORACLE_CLI)
    base = ["-i", f"{pre}.soap", "-d", f"{pre}.fa", "-s", f"{pre}.snp", "-2", "-u", "-L", "100"]
    subprocess.run([cli] + base + ["-o", str(tmp_path / "a.cns")], check=True, capture_output=True)
    subprocess.run([cli] + base + ["-J", "3", "-o", str(tmp_path / "b.cns")], check=True, capture_output=True)
    subprocess.run([cli] + base + ["-J", "2", "-X", "300000", "-o", str(tmp_path / "c.cns")], check=True,
                   capture_output=True)
    a = (tmp_path / "a.cns").read_bytes()
    assert (tmp_path / "b.cns").read_bytes() == a
    # a bounded sample stops at the first window boundary at or past the bound
    assert (tmp_path / "c.cns").read_bytes() == b"".join(a.splitlines(keepends=True)[:2 * 262144])
