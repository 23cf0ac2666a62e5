"""GPU parity: the HIP path against the CPU oracle, byte for byte.

Every test here runs the product through the C ABI on the device and compares
with the oracle on the same seeded inputs: .cns text (plain per-window path and
the batched span path), the -C column files, and per-site unit entry points.
"""
from __future__ import annotations

import ctypes
import filecmp
from pathlib import Path

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

SUFFIXES = [".1_2", ".3_4_6", ".5_7_8_9_14_16", ".10", ".11_12_13", ".15_17"]


def _first_diff(a: Path, b: Path) -> str:
    la = a.read_bytes().split(b"\n")
    lb = b.read_bytes().split(b"\n")
    for i, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {i + 1}:\n gpu    {x!r}\n oracle {y!r}"
    return f"line count {len(la)} vs {len(lb)}"


def _gpu_run(ds, out: Path, **kw):
    import gsnp_amd

    flags = dict(ds.flags)
    flags.update(kw.pop("flags", {}))
    gsnp_amd.run(ds.fasta, ds.soap, out, dbsnp=ds.dbsnp, **kw, **flags)


@pytest.mark.parametrize("name", ["c1", "c2w3", "stress", "c3s", "c5s"])
def test_text_matches_oracle(datasets, tmp_path, name):
    ds = datasets(name)
    print(f"[parity] {name}: {ds.n_reads} reads, flags {ds.flags}", flush=True)
    out = tmp_path / f"{name}.cns"
    _gpu_run(ds, out)
    ref = ds.oracle_text()
    assert out.read_bytes() == ref.read_bytes(), _first_diff(out, ref)


@pytest.mark.parametrize("name,per", [("c2w3", 1), ("c2w3", 8), ("stress", 2), ("c3s", 4), ("c5s", 3)])
def test_batched_matches_oracle(datasets, tmp_path, name, per):
    ds = datasets(name)
    print(f"[parity] batched {name} x{per}", flush=True)
    out = tmp_path / f"{name}.b{per}.cns"
    _gpu_run(ds, out, batched=per)
    ref = ds.oracle_text()
    assert out.read_bytes() == ref.read_bytes(), _first_diff(out, ref)


def _job_text(job) -> bytes:
    import gsnp_amd

    L = gsnp_amd.lib()
    n = ctypes.c_uint64()
    gsnp_amd.check(L.gsnp_fetch_text(job.ctx, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(max(1, n.value))
    gsnp_amd.check(L.gsnp_fetch_text(job.ctx, buf, n.value, ctypes.byref(n)))
    return buf.raw[: n.value]


def _rows_by_contig(ref: bytes) -> dict:
    """The oracle's rows of each contig, in output order.  Lines are gathered
    in lists and joined once: appending to a growing bytes object copies it
    every time, which is quadratic and takes the better part of an hour on the
    600 kb set."""
    rows: dict[bytes, list] = {}
    for line in ref.splitlines(keepends=True):
        rows.setdefault(line.split(b"\t", 1)[0], []).append(line)
    return {k: b"".join(v) for k, v in rows.items()}


@pytest.mark.parametrize("name", ["c2w3", "stress"])
def test_tuned_batched_matches_oracle(datasets, name):
    """The launch shape the autotuner keeps (smallest network width, so small
    sites may run on the 32-lane network, waves per block, grid caps) gives the
    oracle's text: one batched call per contig after tuning."""
    import gsnp_amd

    ds = datasets(name)
    parts = _rows_by_contig(ds.oracle_text().read_bytes())
    job = gsnp_amd.Job(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    try:
        job.select(0)
        print(f"[parity] {name} tuned {job.autotune(reps=1)}", flush=True)
        assert len(parts) == len(job)
        for i in range(len(job)):
            job.select(i)
            want = parts[job.info(i)[0].encode()]
            job.call(batched=64)
            out = _job_text(job)
            assert out == want, f"contig {i}: {len(out)} vs {len(want)} bytes"
    finally:
        job.close()


@pytest.mark.parametrize("name,tune", [("c2w3", False), ("stress", True)])
def test_batched_replays_match_oracle(datasets, name, tune):
    """The batched span runs as a captured graph from its second call on; every
    replay, with eager per-window passes between them (and, with tune, after
    the autotuner has changed the launch shape), gives the oracle's text."""
    import gsnp_amd

    ds = datasets(name)
    ref = ds.oracle_text().read_bytes()
    job = gsnp_amd.Job(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    try:
        job.select(0)
        if tune:
            print(f"[parity] {name} tuned {job.autotune(reps=1)}", flush=True)
        parts = _rows_by_contig(ref)
        assert len(parts) == len(job)
        for i in range(len(job)):
            job.select(i)
            want = parts[job.info(i)[0].encode()]
            for rep in range(4):
                job.call(batched=64)
                out = _job_text(job)
                assert out == want, f"contig {i} replay {rep}: {len(out)} vs {len(want)} bytes"
                job.call(batched=0)  # eager per-window pass between replays
    finally:
        job.close()


@pytest.mark.parametrize("name,flags", [("c1", {"is_monoploid": 1}), ("stress", {"rank_sum_mode": 0}),
                                        ("c2w3", {"pcr_dependency": 1.0, "global_dependency": 1.0}),
                                        ("c2w3", {"pcr_dependency": 0.8, "global_dependency": 0.95,
                                                  "het_novel_r": 0.01})])
@pytest.mark.parametrize("batched", [0, 4])
def test_parameter_variants(datasets, tmp_path, name, flags, batched):
    """Per-window and batched forms under other parameters; the dependency
    coefficients give adjusted-quality tables of 64 bytes (-p 1 -g 1, copied
    into LDS by the fused kernel) and of 85 KB (-p 0.8 -g 0.95, read from
    global memory)."""
    ds = datasets(name)
    print(f"[parity] {name} {flags} batched={batched}", flush=True)
    out = tmp_path / f"{name}.var.cns"
    _gpu_run(ds, out, flags=flags, batched=batched)
    ref = ds.oracle_text(**flags)
    assert out.read_bytes() == ref.read_bytes(), _first_diff(out, ref)


@pytest.mark.parametrize("name,per,lds", [("stress", 0, None), ("c2w3", 0, None), ("stress", 4, None),
                                          ("c2w3", 4, None), ("c2x2", 0, None), ("c2x2", 4, None),
                                          ("c2w3", 4, "0"), ("stress", 0, "0"), ("c2w3", 4, "8"),
                                          ("c5s", 0, None), ("c5s", 3, None), ("c3s", 4, None)])
def test_compressed_output(datasets, tmp_path, monkeypatch, name, per, lds):
    """-C files (encoded on the device, one chunk per reference window; with
    per > 0 the windows of a batched span are encoded together) are the
    oracle's bytes, and decompress to the oracle's text.  By default the
    integer arrays take presence bitmaps and the copy numbers a hash set,
    each array spread over several workgroups; with lds set every array goes
    to the one-workgroup encoder and its hash set, "0" building each one in
    HBM, "8" making most of them overflow the LDS set and start again in
    HBM."""
    import gsnp_amd

    if lds is not None:
        monkeypatch.setenv("GSNP_DICT_LDS_MAX", lds)
        monkeypatch.setenv("GSNP_DICT_BITMAP", "0")
    ds = datasets(name)
    print(f"[parity] -C {name} batched={per} lds={lds}", flush=True)
    gpu = tmp_path / f"{name}.gpu"
    _gpu_run(ds, gpu, compress=True, batched=per)
    cpu = tmp_path / f"{name}.cpu"
    oracle_lib.run(ds.fasta, ds.soap, cpu, dbsnp=ds.dbsnp, compress=True, **ds.flags)
    for sfx in SUFFIXES:
        a, b = Path(str(gpu) + sfx), Path(str(cpu) + sfx)
        assert filecmp.cmp(a, b, shallow=False), f"{sfx} differs"
    # decompressing the GPU files gives the plain text
    txt = tmp_path / f"{name}.dec.cns"
    gsnp_amd.decompress(gpu, txt)
    ref = ds.oracle_text()
    assert txt.read_bytes() == ref.read_bytes(), _first_diff(txt, ref)


@pytest.mark.parametrize("per", [0, 2])
def test_longest_reads_and_capped_sites_match_oracle(built, tmp_path, per):
    """Reads of the largest length (256: cycles up to 255 fill the record's
    cycle field), shorter reads among them, 2- and 3-hit reads, and a pile-up
    of 300 reads at one position, whose sites pass the 255-record cap: the
    calibration, the text (per-window and batched) and the -C files are the
    oracle's."""
    import random

    import gsnp_amd

    rng = random.Random(23)
    G, L = 300000, 256
    ref = "".join(rng.choice("ACGT") for _ in range(G))
    fa, so = tmp_path / "g.fa", tmp_path / "d.soap"
    fa.write_text(">chrL\n" + "\n".join(ref[i:i + 80] for i in range(0, G, 80)) + "\n")
    starts = [rng.randrange(1, G - L - 1) for _ in range(G * 25 // L)] + [150001] * 300
    lines = []
    for i, p in enumerate(sorted(starts)):
        n = L if rng.random() < 0.8 else rng.randint(40, L - 1)
        seq = [b if rng.random() >= 0.01 else rng.choice("ACGT") for b in ref[p - 1:p - 1 + n]]
        q = "".join(chr(64 + rng.randint(2, 40)) for _ in range(n))
        hits = 1 if rng.random() < 0.9 else rng.randint(2, 3)
        lines.append(f"r{i}\t{''.join(seq)}\t{q}\t{hits}\ta\t{n}\t{rng.choice('+-')}\tchrL\t{p}\t0\n")
    so.write_text("".join(lines))
    print(f"[parity] longest reads and capped sites batched={per}", flush=True)
    gt, ct = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    gsnp_amd.run(fa, so, gt, batched=per, read_length=L)
    oracle_lib.run(fa, so, ct, read_length=L)
    assert gt.read_bytes() == ct.read_bytes(), _first_diff(gt, ct)
    deep = [ln for ln in ct.read_text().splitlines() if int(ln.split("\t")[13]) > 255]
    assert deep, "no site past the 255-record cap"
    gpu, cpu = tmp_path / "gpu", tmp_path / "cpu"
    gsnp_amd.run(fa, so, gpu, compress=True, batched=per, read_length=L)
    oracle_lib.run(fa, so, cpu, compress=True, read_length=L)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(gpu) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"


@pytest.mark.parametrize("per", [0, 3])
def test_shared_anchor_reads_match_oracle(built, tmp_path, per):
    """Reads that share their anchor (the start on '+', the last base on '-')
    give a site records of one strand and cycle, which the dependency count
    (kernel/likelihood_kernel.cu:168-256) counts.  chrA has such pairs and
    triples here and there; chrB a block with a pair starting every 4 bp on
    each strand (25 such anchors meet at a site); chrC pairs at every
    position of a block (100 meet).  The text, per-window and batched, is
    the oracle's."""
    import random

    import gsnp_amd

    rng = random.Random(41)
    L, G = 100, 30000
    seqs = {c: "".join(rng.choice("ACGT") for _ in range(G)) for c in ("chrA", "chrB", "chrC")}
    fa, so = tmp_path / "g.fa", tmp_path / "d.soap"
    fa.write_text("".join(f">{k}\n" + "\n".join(v[i:i + 60] for i in range(0, len(v), 60)) + "\n"
                          for k, v in seqs.items()))
    reads = []  # (chromosome, 1-based position, length, strand)

    def scatter(c, n):
        for _ in range(n):
            reads.append((c, rng.randrange(1, G - L - 1), L if rng.random() < 0.8 else rng.randint(30, L), rng.choice("+-")))

    scatter("chrA", G * 20 // L)
    for _ in range(60):  # pairs and triples on one anchor
        p, st, k = rng.randrange(200, G - 300), rng.choice("+-"), rng.choice([2, 2, 3])
        for _ in range(k):
            n = L if st == "+" else rng.randint(60, L)  # '-' reads share their last base
            reads.append(("chrA", p if st == "+" else p + L - n, n, st))
    for c, step in (("chrB", 4), ("chrC", 1)):
        scatter(c, G * 10 // L)
        for a in range(10000, 10400, step):
            for st in "+-":
                for _ in range(2):
                    n = rng.randint(70, L)
                    reads.append((c, a if st == "+" else a + L - n, n, st))
    reads.sort(key=lambda r: (r[0], r[1]))
    # the most anchors held by two or more reads of one strand that meet at a
    # site
    for c, lo, hi in (("chrA", 1, 7), ("chrB", 8, 63), ("chrC", 64, 10 ** 6)):
        most = 0
        for st in "+-":
            held = {}
            for cc, p, n, s2 in reads:
                if cc == c and s2 == st:
                    a = p if st == "+" else p + n - 1
                    held[a] = held.get(a, 0) + 1
            dup = sorted(a for a, k in held.items() if k >= 2)
            j = 0
            for i, a in enumerate(dup):
                while dup[j] <= a - L:
                    j += 1
                most = max(most, i - j + 1)
        assert lo <= most <= hi, (c, most)
    lines = []
    for i, (c, p, n, st) in enumerate(reads):
        ref = seqs[c][p - 1:p - 1 + n]
        seq = "".join(b if rng.random() >= 0.03 else rng.choice("ACGT") for b in ref)
        q = "".join(chr(64 + rng.randint(2, 40)) for _ in range(n))
        lines.append(f"r{i}\t{seq}\t{q}\t1\ta\t{n}\t{st}\t{c}\t{p}\t0\n")
    so.write_text("".join(lines))
    print(f"[parity] shared-anchor reads batched={per}", flush=True)
    gt, ct = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    gsnp_amd.run(fa, so, gt, batched=per, read_length=L)
    oracle_lib.run(fa, so, ct, read_length=L)
    assert gt.read_bytes() == ct.read_bytes(), _first_diff(gt, ct)


@pytest.mark.parametrize("per", [0, 3])
def test_small_empty_and_edge_chromosomes_match_oracle(built, tmp_path, per):
    """Chromosomes of 40 bp (shorter than -L), of 1000 bp without reads, of
    one window plus 10 bp with reads at both ends, and of 500 N: the text
    (per-window and batched) and the -C files are the oracle's."""
    import random

    import gsnp_amd

    rng = random.Random(31)
    W, L = 262144, 50
    seqs = {
        "chrA": "".join(rng.choice("ACGT") for _ in range(40)),
        "chrB": "".join(rng.choice("ACGT") for _ in range(1000)),
        "chrC": "".join(rng.choice("ACGT") for _ in range(W + 10)),
        "chrD": "N" * 500,
    }
    fa, so = tmp_path / "g.fa", tmp_path / "d.soap"
    fa.write_text("".join(f">{k}\n" + "\n".join(v[i:i + 60] for i in range(0, len(v), 60)) + "\n"
                          for k, v in seqs.items()))
    reads = []  # (chromosome, 1-based position, length)
    reads += [("chrA", p, 30) for p in (1, 1, 5, 11, 11)]
    C = len(seqs["chrC"])
    reads += [("chrC", p, L) for p in sorted([1, 1, 2] + [rng.randrange(1, C - L + 2) for _ in range(W * 8 // L)]
                                             + [C - L + 1, C - L + 1, C - L])]
    lines = []
    for i, (c, p, n) in enumerate(reads):
        ref = seqs[c][p - 1:p - 1 + n]
        seq = "".join(b if rng.random() >= 0.02 else rng.choice("ACGT") for b in ref)
        q = "".join(chr(64 + rng.randint(2, 40)) for _ in range(n))
        lines.append(f"r{i}\t{seq}\t{q}\t1\ta\t{n}\t{rng.choice('+-')}\t{c}\t{p}\t0\n")
    so.write_text("".join(lines))
    print(f"[parity] small, empty and edge chromosomes batched={per}", flush=True)
    gt, ct = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    gsnp_amd.run(fa, so, gt, batched=per, read_length=L)
    oracle_lib.run(fa, so, ct, read_length=L)
    assert gt.read_bytes() == ct.read_bytes(), _first_diff(gt, ct)
    gpu, cpu = tmp_path / "gpu", tmp_path / "cpu"
    gsnp_amd.run(fa, so, gpu, compress=True, batched=per, read_length=L)
    oracle_lib.run(fa, so, cpu, compress=True, read_length=L)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(gpu) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"


@pytest.mark.parametrize("per", [0, 4])
def test_compressed_many_distinct_values(built, tmp_path, per):
    """Reads with random hit counts give a window tens of thousands of
    distinct copy numbers (column 16): the dictionary's hash set outgrows LDS
    and restarts in HBM, and the ranks take many tiles.  The -C files are the
    oracle's bytes."""
    import random

    import gsnp_amd

    rng = random.Random(7)
    G, L = 300000, 50
    ref = "".join(rng.choice("ACGT") for _ in range(G))
    fa, so = tmp_path / "g.fa", tmp_path / "d.soap"
    fa.write_text(">chrD\n" + "\n".join(ref[i:i + 80] for i in range(0, G, 80)) + "\n")
    lines = []
    for i, p in enumerate(sorted(rng.randrange(1, G - L - 1) for _ in range(G * 20 // L))):
        seq = [b if rng.random() >= 0.01 else rng.choice("ACGT") for b in ref[p - 1:p - 1 + L]]
        q = "".join(chr(64 + rng.randint(5, 40)) for _ in range(L))
        hits = 1 if rng.random() < 0.4 else rng.randint(2, 400)
        lines.append(f"r{i}\t{''.join(seq)}\t{q}\t{hits}\ta\t{L}\t{rng.choice('+-')}\tchrD\t{p}\t0\n")
    so.write_text("".join(lines))
    print(f"[parity] -C many distinct values batched={per}", flush=True)
    gpu, cpu = tmp_path / "gpu", tmp_path / "cpu"
    gsnp_amd.run(fa, so, gpu, compress=True, batched=per, read_length=L)
    oracle_lib.run(fa, so, cpu, compress=True, read_length=L)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(gpu) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"
    # the same calls as text: thousands of distinct %#g copy numbers
    gt, ct = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    gsnp_amd.run(fa, so, gt, batched=per, read_length=L)
    oracle_lib.run(fa, so, ct, read_length=L)
    assert gt.read_bytes() == ct.read_bytes(), _first_diff(gt, ct)


@pytest.mark.parametrize("per", [0, 4])
def test_compressed_sorted_copy_number_ranks(built, tmp_path, per):
    """Windows with a few thousand distinct copy numbers (column 16): about
    3,400 on chrA and 5,900 on chrB, above the 1,024 the dictionary ranks by
    counting and within the 8,192 it ranks by a sort in LDS.  The -C files and
    the text are the oracle's."""
    import random

    import gsnp_amd

    rng = random.Random(11)
    G, L = 100000, 50
    seqs = {c: "".join(rng.choice("ACGT") for _ in range(G)) for c in ("chrA", "chrB")}
    fa, so = tmp_path / "g.fa", tmp_path / "d.soap"
    fa.write_text("".join(f">{k}\n" + "\n".join(v[i:i + 80] for i in range(0, G, 80)) + "\n"
                          for k, v in seqs.items()))
    lines = []
    for c, hmax in (("chrA", 30), ("chrB", 60)):
        starts = sorted(rng.randrange(1, G - L - 1) for _ in range(G * 20 // L))
        hits = [1 if rng.random() < 0.5 else rng.randint(2, hmax) for _ in starts]
        # the window's distinct copy numbers, repeat / depth over covered sites
        depth, rep = [0] * (G + 1), [0] * (G + 1)
        for p, h in zip(starts, hits):
            depth[p - 1] += 1
            depth[p - 1 + L] -= 1
            rep[p - 1] += h
            rep[p - 1 + L] -= h
        vals, dd, rr = set(), 0, 0
        for x in range(G):
            dd += depth[x]
            rr += rep[x]
            if dd:
                vals.add(rr / dd)
        assert 1024 < len(vals) + 1 <= 8192, (c, len(vals))
        for p, h in zip(starts, hits):
            ref = seqs[c][p - 1:p - 1 + L]
            seq = "".join(b if rng.random() >= 0.01 else rng.choice("ACGT") for b in ref)
            q = "".join(chr(64 + rng.randint(5, 40)) for _ in range(L))
            lines.append(f"r{len(lines)}\t{seq}\t{q}\t{h}\ta\t{L}\t{rng.choice('+-')}\t{c}\t{p}\t0\n")
    so.write_text("".join(lines))
    print(f"[parity] -C sorted copy-number ranks batched={per}", flush=True)
    gpu, cpu = tmp_path / "gpu", tmp_path / "cpu"
    gsnp_amd.run(fa, so, gpu, compress=True, batched=per, read_length=L)
    oracle_lib.run(fa, so, cpu, compress=True, read_length=L)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(gpu) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"
    gt, ct = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    gsnp_amd.run(fa, so, gt, batched=per, read_length=L)
    oracle_lib.run(fa, so, ct, read_length=L)
    assert gt.read_bytes() == ct.read_bytes(), _first_diff(gt, ct)


@pytest.mark.parametrize("per", [0, 4])
def test_compressed_uncovered_window(built, tmp_path, per):
    """A reference window without reads: its run-length columns hold one run
    of 262,144 sites, a value beyond the integer dictionaries' presence bitmap
    ([0, 262144)), so those arrays go back to the one-workgroup encoder while
    the other windows' arrays take the split bitmap path.  The -C files are
    the oracle's bytes."""
    import random

    import gsnp_amd

    rng = random.Random(11)
    W, L = 262144, 50
    G = 2 * W + 150000
    ref = "".join(rng.choice("ACGT") for _ in range(G))
    fa, so = tmp_path / "g.fa", tmp_path / "d.soap"
    fa.write_text(">chrU\n" + "\n".join(ref[i:i + 80] for i in range(0, G, 80)) + "\n")
    starts = sorted(p for p in (rng.randrange(1, G - L - 1) for _ in range(G * 6 // L))
                    if p + L <= W or p > 2 * W)  # nothing reaches window 1
    lines = []
    for i, p in enumerate(starts):
        seq = [b if rng.random() >= 0.01 else rng.choice("ACGT") for b in ref[p - 1:p - 1 + L]]
        q = "".join(chr(64 + rng.randint(5, 40)) for _ in range(L))
        lines.append(f"r{i}\t{''.join(seq)}\t{q}\t1\ta\t{L}\t{rng.choice('+-')}\tchrU\t{p}\t0\n")
    so.write_text("".join(lines))
    print(f"[parity] -C uncovered window batched={per}", flush=True)
    gpu, cpu = tmp_path / "gpu", tmp_path / "cpu"
    gsnp_amd.run(fa, so, gpu, compress=True, batched=per, read_length=L)
    oracle_lib.run(fa, so, cpu, compress=True, read_length=L)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(gpu) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"
    dec = tmp_path / "dec.cns"
    gsnp_amd.decompress(gpu, dec)
    ct = tmp_path / "cpu.cns"
    oracle_lib.run(fa, so, ct, read_length=L)
    assert dec.read_bytes() == ct.read_bytes(), _first_diff(dec, ct)


def test_compressed_long_span(built, tmp_path):
    """A 9 Mb chromosome (35 windows, more than one encoder pass of 32 windows):
    the -C files of one batched span equal those of the per-window path, and
    decompress to the batched span's text."""
    import gsnp_amd

    pre = tmp_path / "c2h"
    _, rl = gsnp_amd.synth(2, 7, 0.9, pre)
    fa, so, snp = f"{pre}.fa", f"{pre}.soap", f"{pre}.snp"
    flags = dict(read_length=rl, refine_mode=1)
    print("[parity] -C long span", flush=True)
    a, b, txt = tmp_path / "a", tmp_path / "b", tmp_path / "t.cns"
    gsnp_amd.run(fa, so, a, dbsnp=snp, compress=True, batched=0, **flags)
    gsnp_amd.run(fa, so, b, dbsnp=snp, compress=True, batched=64, **flags)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(a) + sfx, str(b) + sfx, shallow=False), f"{sfx} differs"
    gsnp_amd.run(fa, so, txt, dbsnp=snp, batched=64, **flags)
    dec = tmp_path / "b.dec.cns"
    gsnp_amd.decompress(b, dec)
    assert dec.read_bytes() == txt.read_bytes(), _first_diff(dec, txt)


@pytest.mark.parametrize("name", ["c2w3", "stress", "c3s"])
def test_calibration_matches_oracle(datasets, name):
    """Pass-1 calibration counted on the device (gsnp_calib_count over every
    chromosome's reads, then the matrix on the host) equals the oracle's
    matrix_gen, double for double; c3s has 24 chromosomes, stress known SNPs."""
    import gsnp_amd

    ds = datasets(name)
    print(f"[parity] calibration {name}", flush=True)
    gpu = gsnp_amd.matrix_gen(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    ref = oracle_lib.matrix_gen(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    bad = np.flatnonzero(gpu.view(np.uint64) != ref.view(np.uint64))
    assert bad.size == 0, f"{bad.size} entries differ, first at {bad[0]}: {gpu[bad[0]]!r} vs {ref[bad[0]]!r}"


def test_sort_likelihood_unit(built):
    """Random record sets through the sort + likelihood kernels vs the oracle."""
    import gsnp_amd

    L = gsnp_amd.lib()
    o = oracle_lib.orc()
    rng = np.random.default_rng(5)
    p = gsnp_amd.params(read_length=100)
    pm = np.full(256 * 256 * 16, 1.0)
    # a plausible calibration: match 1-e, mismatch e/3 by quality
    q = np.arange(256)
    e = 10.0 ** (-q / 10.0)
    for t in range(4):
        for ob in range(4):
            col = np.where(t == ob, 1 - e, e / 3)
            idx = (q[:, None] << 12) | (np.arange(256)[None, :] << 4) | (t << 2) | ob
            pm[idx] = col[:, None]
    ctx = ctypes.c_void_p()
    gsnp_amd.check(L.gsnp_device_init(0))
    gsnp_amd.check(L.gsnp_create(ctypes.byref(p), pm.ctypes.data, ctypes.byref(ctx)))
    try:
        n_sites = 3000
        counts = rng.integers(0, 256, n_sites).astype(np.uint32)
        counts[:40] = [0, 1, 2, 8, 9, 16, 17, 32, 33, 64, 65, 128, 129, 255, 256, 7, 15, 31, 63, 127] * 2
        recs = np.zeros((n_sites, 256), dtype=np.uint32)
        for i in range(n_sites):
            n = counts[i]
            base = rng.integers(0, 4, n)
            qs = rng.integers(1, 41, n)
            cyc = rng.integers(0, 100, n)
            st = rng.integers(0, 2, n)
            if i % 3 == 0:  # force duplicate (strand, cycle) keys
                cyc = rng.integers(0, 4, n)
            recs[i, :n] = (base << 23) | ((64 - qs) << 17) | (cyc << 9) | (st << 8) | 1
        lk = np.zeros((n_sites, 10))
        work = recs.copy()
        gsnp_amd.check(L.gsnp_sort_likelihood(ctx, work.ctypes.data, 256, counts.ctypes.data, n_sites, lk.ctypes.data))
        score = np.zeros(256 * 256 * 4 * 10)
        o.orc_score_table(pm.ctypes.data, score.ctypes.data)
        for i in range(n_sites):
            n = int(counts[i])
            want = np.sort(recs[i, :n])
            assert np.array_equal(work[i, :n], want), f"site {i} (n={n}) not sorted"
            ref = np.zeros(10)
            rr = np.ascontiguousarray(want)
            assert o.orc_likelihood_sorted(rr.ctypes.data, n, 100, p.pcr_dependency, p.global_dependency,
                                           score.ctypes.data, ref.ctypes.data) == 0
            assert np.array_equal(lk[i].view(np.uint64), ref.view(np.uint64)), f"site {i} likelihood differs"
    finally:
        L.gsnp_destroy(ctx)


def _unpack(words, n, nd):
    """Ranks from MSB-first packed words (cpp/dict_compression.h:116-166):
    floor(32 / bits) per word, the last word right-aligned."""
    bits = max(1, int(np.ceil(np.log(float(nd)) / np.log(2.0)))) if nd > 1 else 1
    per = 32 // bits
    out = []
    nw = (n + per - 1) // per
    for i in range(nw):
        m = n - per * i if i == nw - 1 else per
        for k in range(m):
            out.append((int(words[i]) >> (bits * (m - k - 1))) & ((1 << bits) - 1))
    return np.array(out, np.int64)


@pytest.mark.parametrize("kind", ["char", "u32", "f64"])
def test_device_dictionary_overloads_match_reference_codec(built, kind):
    """The char, uint32 and double overloads of cuda_dictCompression
    (cpp/dict_compression.h:26-37) against the reference's own template
    compiled from its header, and against what the format fixes: the
    dictionary is the sorted distinct values and the ranks index it."""
    import gsnp_amd

    L = gsnp_amd.lib()
    ref = oracle_lib.dictref()
    p = gsnp_amd.params(read_length=100)
    pm = np.full(256 * 256 * 16, 0.5)
    ctx = ctypes.c_void_p()
    gsnp_amd.check(L.gsnp_device_init(0))
    gsnp_amd.check(L.gsnp_create(ctypes.byref(p), pm.ctypes.data, ctypes.byref(ctx)))
    rng = np.random.default_rng(11)
    dt = {"char": np.int8, "u32": np.uint32, "f64": np.float64}[kind]
    if kind == "char":
        cases = [rng.integers(-128, 128, n).astype(np.int8) for n in (1, 5, 300, 70000)]
        cases.append(np.frombuffer(b"ACGTNACGTTTN" * 50, np.int8).copy())
    elif kind == "u32":
        cases = [rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32) for n in (1, 7, 3000)]
        cases += [rng.integers(0, 300, 70000).astype(np.uint32), np.array([2**32 - 1, 0, 2**31, 2**31 - 1] * 9, np.uint32)]
    else:
        cases = [np.round(rng.normal(0, 50, n), 1) for n in (1, 9, 5000)]
        cases += [np.array([1.0, 255.0, 0.5, -3.25, 1e-300, -1e300, np.inf, -np.inf, 0.0] * 7),
                  rng.choice(np.array([0.125, 1.0, 2.5, 1e9]), 40000), np.array([-0.0, 1.0, -0.0])]
    fn = getattr(L, f"gsnp_dict_compress_{kind}")
    try:
        for v in cases:
            v = np.ascontiguousarray(v.astype(dt))
            n = len(v)
            w = np.zeros(n + 1, np.uint32)
            d = np.zeros(n + 1, dt)
            nw, nd = ctypes.c_uint64(), ctypes.c_uint64()
            gsnp_amd.check(fn(ctx, v.ctypes.data, n, w.ctypes.data, ctypes.byref(nw), d.ctypes.data, ctypes.byref(nd)))
            dict_ = d[:nd.value]
            assert np.array_equal(dict_, np.unique(v)), "dictionary is not the sorted distinct values"
            assert np.array_equal(dict_[_unpack(w, n, nd.value)], v), "ranks do not index the values"
            zs = np.signbit(v[v == 0]) if kind == "f64" else np.zeros(0, bool)
            if zs.size and (zs.all() or not zs.any()):  # every zero of one sign: the dictionary keeps it
                assert np.signbit(dict_[dict_ == 0]).tolist() == [bool(zs[0])]
            if ref is not None:
                w3 = np.zeros(n + 1, np.uint32)
                d3 = np.zeros(n + 1, dt)
                nd3 = ctypes.c_ulong()
                nw3 = getattr(ref, f"ref_dict_encode_{kind}")(v.ctypes.data, n, w3.ctypes.data, d3.ctypes.data,
                                                              ctypes.byref(nd3))
                assert (nw3, nd3.value) == (nw.value, nd.value)
                assert np.array_equal(w[:nw3], w3[:nw3])
                assert d[:nd3.value].tobytes() == d3[:nd3.value].tobytes()
        nan = np.array([1.0, np.nan])
        if kind == "f64":
            w = np.zeros(3, np.uint32)
            d = np.zeros(3)
            nw, nd = ctypes.c_uint64(), ctypes.c_uint64()
            assert fn(ctx, nan.ctypes.data, 2, w.ctypes.data, ctypes.byref(nw), d.ctypes.data, ctypes.byref(nd)) != 0
    finally:
        L.gsnp_destroy(ctx)


def test_device_dictionary_matches_reference_codec(built):
    """gsnp_dict_compress_i32 vs the reference's own dictCompression (and the oracle)."""
    import gsnp_amd

    L = gsnp_amd.lib()
    ref = oracle_lib.dictref()
    o = oracle_lib.orc()
    p = gsnp_amd.params(read_length=100)
    pm = np.full(256 * 256 * 16, 0.5)
    ctx = ctypes.c_void_p()
    gsnp_amd.check(L.gsnp_device_init(0))
    gsnp_amd.check(L.gsnp_create(ctypes.byref(p), pm.ctypes.data, ctypes.byref(ctx)))
    rng = np.random.default_rng(9)
    try:
        cases = [np.array([7], np.int32), np.array([3, 1, 3, 7, 1], np.int32), np.arange(1024, dtype=np.int32),
                 rng.integers(0, 2, 5000).astype(np.int32), rng.integers(-50, 50, 10001).astype(np.int32),
                 rng.integers(0, 65537, 70000).astype(np.int32), rng.integers(-2**31, 2**31 - 1, 3000).astype(np.int32)]
        for k in [1, 2, 3, 4, 5, 8, 9, 16, 17, 255, 256, 257, 1023, 1024, 1025]:
            cases.append(rng.permutation(np.repeat(np.arange(k, dtype=np.int32) * 3 - 7, 3)))
        for v in cases:
            n = len(v)
            w = np.zeros(n + 1, np.uint32)
            d = np.zeros(n + 1, np.int32)
            nw, nd = ctypes.c_uint64(), ctypes.c_uint64()
            gsnp_amd.check(L.gsnp_dict_compress_i32(ctx, v.ctypes.data, n, w.ctypes.data, ctypes.byref(nw),
                                                    d.ctypes.data, ctypes.byref(nd)))
            w2 = np.zeros(n + 1, np.uint32)
            d2 = np.zeros(n + 1, np.int32)
            nd2 = ctypes.c_uint64()
            nw2 = o.orc_dict_encode_i32(v.ctypes.data, n, w2.ctypes.data, d2.ctypes.data, ctypes.byref(nd2))
            assert (nw.value, nd.value) == (nw2, nd2.value)
            assert np.array_equal(w[:nw2], w2[:nw2]) and np.array_equal(d[:nd2.value], d2[:nd2.value])
            if ref is not None:
                w3 = np.zeros(n + 1, np.uint32)
                d3 = np.zeros(n + 1, np.int32)
                nd3 = ctypes.c_ulong()
                nw3 = ref.ref_dict_encode_i32(v.ctypes.data, n, w3.ctypes.data, d3.ctypes.data, ctypes.byref(nd3))
                assert (nw3, nd3.value) == (nw.value, nd.value)
                assert np.array_equal(w[:nw3], w3[:nw3]) and np.array_equal(d[:nd3.value], d3[:nd3.value])
    finally:
        L.gsnp_destroy(ctx)


@pytest.mark.parametrize("case", ["sort_then_quality", "quality_then_sort", "length"])
def test_load_reads_reports_the_first_read_error(built, case):
    """gsnp_load_reads validates and packs ranges of reads on threads: the
    error reported is the one of the lowest read index, as a read-by-read
    loop finds it."""
    import gsnp_amd

    L = gsnp_amd.lib()
    n, rl = 400000, 10
    p = gsnp_amd.params(read_length=rl)
    pm = np.full(256 * 256 * 16, 0.25)
    ctx = ctypes.c_void_p()
    gsnp_amd.check(L.gsnp_device_init(0))
    gsnp_amd.check(L.gsnp_create(ctypes.byref(p), pm.ctypes.data, ctypes.byref(ctx)))
    try:
        G = n + 100
        codes = np.zeros(G, dtype=np.uint8)
        gsnp_amd.check(L.gsnp_set_chromosome(ctx, b"chr1", codes.ctypes.data, G, None, None, 0))
        pos = np.arange(n, dtype=np.int32)
        len_hit = np.full(n, (rl << 16) | 1, dtype=np.int32)
        strand = np.full(n, ord("+"), dtype=np.uint8)
        seq = np.full(n * rl, ord("A"), dtype=np.uint8)
        qual = np.full(n * rl, ord("I"), dtype=np.uint8)
        if case == "sort_then_quality":
            pos[100000] = 5
            qual[300000 * rl + 3] = ord("@")  # q - q_min = 0 on a unique-hit base
            want = (3, b"Errors in sorting:5<99999")
        elif case == "quality_then_sort":
            qual[100000 * rl + 3] = ord("@")
            pos[300000] = 5
            want = (5, b"quality outside [1,63] on a unique-hit base")
        else:
            len_hit[350000] = ((rl + 1) << 16) | 1
            qual[360000 * rl] = ord("@")
            want = (6, b"read length outside [1, read_length]")
        b = gsnp_amd.ReadBatch(n, rl, pos.ctypes.data, len_hit.ctypes.data, strand.ctypes.data, seq.ctypes.data,
                               qual.ctypes.data)
        rc = L.gsnp_load_reads(ctx, ctypes.byref(b))
        assert (rc, L.gsnp_last_error()) == want
    finally:
        L.gsnp_destroy(ctx)


def test_full_size_batched_equals_per_window(built, tmp_path):
    """The 10 Mb workload (30x, 3 M reads, 291 M records): one batched span of
    39 windows writes the same .cns bytes as the per-window path (a
    size-independent check at the benchmark's full size, where the oracle is
    too slow to run)."""
    import gsnp_amd

    pre = tmp_path / "c2"
    _, rl = gsnp_amd.synth(2, 2024, 1.0, pre)
    fa, so, snp = f"{pre}.fa", f"{pre}.soap", f"{pre}.snp"
    print("[parity] 10 Mb batched vs per-window", flush=True)
    a, b = tmp_path / "a.cns", tmp_path / "b.cns"
    gsnp_amd.run(fa, so, a, dbsnp=snp, batched=64, read_length=rl, refine_mode=1)
    gsnp_amd.run(fa, so, b, dbsnp=snp, batched=0, read_length=rl, refine_mode=1)
    assert a.stat().st_size > 500_000_000
    assert filecmp.cmp(a, b, shallow=False), _first_diff(a, b)


@pytest.mark.parametrize("name,per", [("c2w3", 2), ("c3s", 0), ("c5s", 64)])
def test_synthetic_job_matches_oracle(datasets, tmp_path, name, per):
    """A job over the data set generated in memory (gsnp_job_synth, the
    benchmark's input path) writes the oracle's text for the files the same
    (config, seed, scale) generates."""
    import gsnp_amd
    from conftest import DATASETS

    ds = datasets(name)
    cfg, seed, scale, _ = DATASETS[name]
    flags = {k: v for k, v in ds.flags.items() if k != "read_length"}
    print(f"[parity] in-memory {name} batched={per}", flush=True)
    job = gsnp_amd.Job(synth=(cfg, seed, scale), **flags)
    try:
        assert job.read_len == ds.read_len
        out = tmp_path / f"{name}.mem.cns"
        job.write(out, batched=per)
        ref = ds.oracle_text()
        assert out.read_bytes() == ref.read_bytes(), _first_diff(out, ref)
    finally:
        job.close()


@pytest.mark.parametrize("name", ["c3s", "c5s"])
def test_whole_pass_rows_reach_host_memory(datasets, name):
    """gsnp_job_pass over every chromosome: the rows copied to host memory
    (batched spans through the overlapped copy, and the per-window path with
    one copy per window) add up to the oracle's text, and the rank-sum sites
    finished on the host are counted (printed for the record)."""
    import gsnp_amd

    ds = datasets(name)
    size = ds.oracle_text().stat().st_size
    job = gsnp_amd.Job(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    try:
        sites = sum(job.info(i)[1] for i in range(len(job)))
        job.counters(reset=True)
        for per, mode in ((-1, 2), (4, 2), (0, 2), (-1, 1)):
            nb, ns = job.run_pass(batched=per, emit_text=mode)
            assert (nb, ns) == (size, sites), (per, mode, nb, ns)
            c = job.counters(reset=True)
            print(f"[parity] {name} pass batched={per} text={mode}: {c}", flush=True)
            assert c["records"] > 0 and c["host_fixups"] >= c["host_fixups_normal"]
    finally:
        job.close()


@pytest.mark.parametrize("name", ["stress", "c3s"])
@pytest.mark.parametrize("per", [0, 4])
def test_rank_sum_normal_approximation_on_device(datasets, tmp_path, name, per):
    """Het sites with n1 + n2 >= 64 (stress: about 490, c3s: about 40, in its
    repeat-like blocks) take the rank-sum test's normal approximation; the
    device finishes it, with exp and sqrt giving the host libm's doubles,
    except where a value lies within 2 % of an ulp of a rounding boundary.
    The text is the oracle's, and the counters show both paths."""
    import gsnp_amd

    ds = datasets(name)
    ref = ds.oracle_text()
    job = gsnp_amd.Job(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    try:
        job.counters(reset=True)
        out = tmp_path / f"{name}.n{per}.cns"
        job.write(out, batched=per)
        c = job.counters(reset=True)
        print(f"[parity] {name} batched={per} rank-sum counters {c}", flush=True)
        assert out.read_bytes() == ref.read_bytes(), _first_diff(out, ref)
        assert c["device_normal"] > 0
        assert c["host_fixups_normal"] < c["device_normal"]
    finally:
        job.close()


REGIONS = {
    # c2w3 (one 600 kb chromosome, windows 0-2): a region inside window 0, one
    # whose read-length extension crosses into window 1 (window 2 is not
    # touched), and a line of an unknown chromosome (skipped)
    "c2w3": "chr1\t100\t200\nchr1\t262100\t262200\nchrZ\t1\t10\n",
    # stress (chr1 560 kb, chr2 300 kb, chr3 90 kb): window 1 of chr1 only,
    # window 0 of chr3 (its start clipped at 0), nothing on chr2
    "stress": "chr1\t300000\t300050\nchr3\t10\t20\n",
}


@pytest.mark.parametrize("name,per,compress", [("c2w3", 0, False), ("c2w3", 4, False), ("stress", 0, False),
                                               ("stress", 2, False), ("c2w3", 0, True), ("stress", 2, True)])
def test_regions_match_oracle(datasets, tmp_path, name, per, compress):
    """-T target regions: only the windows a region touches are called and
    written (text or -C chunks), each from every read that covers it."""
    import gsnp_amd

    ds = datasets(name)
    reg = tmp_path / "regions.txt"
    reg.write_text(REGIONS[name])
    print(f"[parity] -T {name} batched={per} compress={compress}", flush=True)
    gpu, cpu = tmp_path / "gpu", tmp_path / "cpu"
    job = gsnp_amd.Job(ds.fasta, ds.soap, dbsnp=ds.dbsnp, **ds.flags)
    try:
        job.set_regions(reg)
        job.write(gpu, compress=compress, batched=per)
    finally:
        job.close()
    oracle_lib.run(ds.fasta, ds.soap, cpu, dbsnp=ds.dbsnp, compress=compress, regions=reg, **ds.flags)
    if compress:
        for sfx in SUFFIXES:
            assert filecmp.cmp(str(gpu) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"
        assert Path(str(gpu) + ".1_2").read_text().splitlines()[0] == "2"  # two windows touched
    else:
        a, b = gpu.read_bytes(), cpu.read_bytes()
        assert a == b, _first_diff(gpu, cpu)
        # the called windows are whole: every row belongs to a touched window
        full = ds.oracle_text().read_bytes().splitlines(keepends=True)
        assert 0 < len(a.splitlines()) < len(full)


@pytest.mark.parametrize("name", ["stress", "c2x2", "c5s", "c3s"])
def test_device_decompression_matches_oracle(datasets, tmp_path, name):
    """-C files decompressed on the device (gsnp_decompress_device: columns
    rebuilt and rows written by kernels) give the oracle's text; c2x2 ends
    with an empty window, stress has all-N windows, known SNPs and rank sums."""
    import gsnp_amd

    ds = datasets(name)
    pre = tmp_path / name
    oracle_lib.run(ds.fasta, ds.soap, pre, dbsnp=ds.dbsnp, compress=True, **ds.flags)
    print(f"[parity] device decompression {name}", flush=True)
    out = tmp_path / f"{name}.dev.cns"
    gsnp_amd.decompress(pre, out, device=True)
    ref = ds.oracle_text()
    assert out.read_bytes() == ref.read_bytes(), _first_diff(out, ref)


def _assemble_cfiles(prefix: Path, windows) -> None:
    """-C files from (chromosome, start, sites, five chunks) per window, in
    the layout of cpp/compress.h: the chunks appended file by file, and the
    .1_2 index (window count, then one chromosome/start/size row each)."""
    with open(str(prefix) + ".1_2", "w") as f:
        f.write(f"{len(windows)}\n")
        for name, start, n, _ in windows:
            f.write(f"{name}\t{start}\t{n}\n")
    for k, sfx in enumerate(SUFFIXES[1:]):
        with open(str(prefix) + sfx, "wb") as f:
            for *_, chunks in windows:
                f.write(chunks[k])


@pytest.mark.parametrize("name", ["stress", "c2x2", "c5s"])
def test_window_and_callback_columns_match_oracle(datasets, tmp_path, name):
    """The -C chunks through the two chunk-level entry points: the window
    driver (init, update, likelihood, posterior, gsnp_window_columns,
    recycle per window, the structure of cpp/call_genotype.cc:474-644 with
    the -C branch of call_cns2) and the whole-chromosome callback
    (gsnp_job_columns_to, batched spans of 3 windows); both assembled into
    files equal the oracle's -C files byte for byte.  c2x2 ends with an empty
    window."""
    import gsnp_amd

    ds = datasets(name)
    cpu = tmp_path / f"{name}.cpu"
    oracle_lib.run(ds.fasta, ds.soap, cpu, dbsnp=ds.dbsnp, compress=True, **ds.flags)
    job = gsnp_amd.Job(ds.fasta, ds.soap, ds.dbsnp, **ds.flags)
    L = gsnp_amd.lib()
    W = 262144
    per_window, callback = [], []
    for i in range(len(job)):
        cname, N, n_reads = job.info(i)
        print(f"[parity] window columns {name} {cname}", flush=True)
        job.select(i)
        ctx = job.ctx
        gsnp_amd.check(L.gsnp_window_init(ctx, 0))
        for w in range(N // W + 1):
            C = W if w < N // W else N % W
            # every read of the batch: the update keeps those starting in the window
            gsnp_amd.check(L.gsnp_window_update(ctx, 0, n_reads, None))
            gsnp_amd.check(L.gsnp_window_likelihood(ctx, C))
            gsnp_amd.check(L.gsnp_window_posterior(ctx, C))
            per_window.append((cname, w * W + 1, C, gsnp_amd.window_columns(ctx, C)))
            gsnp_amd.check(L.gsnp_window_recycle(ctx))
        callback += job.column_chunks(batched=3)
    for label, windows in (("window", per_window), ("callback", callback)):
        pre = tmp_path / f"{name}.{label}"
        _assemble_cfiles(pre, windows)
        for sfx in SUFFIXES:
            a, b = Path(str(pre) + sfx), Path(str(cpu) + sfx)
            assert filecmp.cmp(a, b, shallow=False), f"{label}: {sfx} differs"


def test_chunk_callback_error_leaves_no_queued_pass(datasets, tmp_path):
    """A chunk callback that fails on its second chunk ends gsnp_job_columns_to
    with an I/O error while a later pass is still queued (spans of one
    window); that pass is dropped, so a full callback pass on the same job
    afterwards gives the oracle's -C files, and gsnp_job_pass with -C chunks
    counts the same bytes."""
    import gsnp_amd

    ds = datasets("c2w3")
    cpu = tmp_path / "c2w3.cpu"
    oracle_lib.run(ds.fasta, ds.soap, cpu, dbsnp=ds.dbsnp, compress=True, **ds.flags)
    job = gsnp_amd.Job(ds.fasta, ds.soap, ds.dbsnp, **ds.flags)
    L = gsnp_amd.lib()
    seen = []

    def fail_second(user, chr_, start, n, chunk, nbytes):
        seen.append(start)
        return 1 if len(seen) == 2 else 0

    print("[parity] chunk callback error, then a full pass", flush=True)
    windows = []
    for i in range(len(job)):
        job.select(i)
        if i == 0:
            cb = gsnp_amd.CHUNK_FN(fail_second)
            nb = ctypes.c_uint64()
            rc = L.gsnp_job_columns_to(job.h, 1, cb, None, ctypes.byref(nb))
            assert rc == 1 and len(seen) == 2, (rc, seen)  # GSNP_E_IO
        windows += job.column_chunks(batched=1)
    pre = tmp_path / "c2w3.after"
    _assemble_cfiles(pre, windows)
    for sfx in SUFFIXES:
        assert filecmp.cmp(str(pre) + sfx, str(cpu) + sfx, shallow=False), f"{sfx} differs"
    nbytes, nsites = job.run_pass(batched=1, emit_text=3)
    assert nbytes == sum(len(c) for w in windows for c in w[3])


def test_caller_stream_gives_the_same_rows(datasets, tmp_path):
    """gsnp_set_stream queues the context's work on a caller's stream (one
    created through the HIP runtime the library uses, as a C caller would);
    the rows written through it, and after switching back to the context's
    own stream, are the oracle's."""
    import gsnp_amd

    ds = datasets("c2w3")
    job = gsnp_amd.Job(ds.fasta, ds.soap, ds.dbsnp, **ds.flags)
    L = gsnp_amd.lib()
    hip = ctypes.CDLL("libamdhip64.so.7")  # already loaded by the library: the same runtime
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    ref = ds.oracle_text().read_bytes()
    try:
        print("[parity] caller stream c2w3", flush=True)
        gsnp_amd.check(L.gsnp_set_stream(job.ctx, stream))
        out = tmp_path / "caller.cns"
        job.write(out, batched=4)
        assert out.read_bytes() == ref, _first_diff(out, ds.oracle_text())
        gsnp_amd.check(L.gsnp_set_stream(job.ctx, None))
        out2 = tmp_path / "own.cns"
        job.write(out2, batched=0)
        assert out2.read_bytes() == ref, _first_diff(out2, ds.oracle_text())
    finally:
        job.close()
        assert hip.hipStreamDestroy(stream) == 0
