"""The seeded test data sets hold what the tests built on them rely on (CPU only).

c3s, the small form of the uneven-depth configuration, must contain
repeat-like blocks (depth 200 and more with 2- and 3-hit reads) and uncovered
blocks; c5s has 150 bp reads; c2x2 ends exactly on a window boundary.  The
in-memory generator is the file generator: the same triple gives the same
inputs whatever the thread count."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest


def _soap_hits(path: Path) -> dict:
    hits: dict[int, int] = {}
    with open(path) as f:
        for line in f:
            h = int(line.split("\t", 4)[3])
            hits[h] = hits.get(h, 0) + 1
    return hits


def test_c3s_has_repeat_and_uncovered_blocks(datasets):
    ds = datasets("c3s")
    hits = _soap_hits(ds.soap)
    assert hits.get(2, 0) > 0 and hits.get(3, 0) > 0, hits
    deep = uncovered = 0
    with open(ds.oracle_text()) as f:
        for line in f:
            d = int(line.split("\t")[13])
            deep += d >= 200
            uncovered += d == 0
    print(f"c3s: {deep} sites at depth >= 200, {uncovered} uncovered, hits {hits}")
    assert deep >= 5000, "no repeat-like block at depth 200-255"
    assert uncovered >= 5000, "no uncovered block"


def test_c5s_reads_are_150bp(datasets):
    ds = datasets("c5s")
    assert ds.read_len == 150
    with open(ds.soap) as f:
        assert len(f.readline().split("\t")[1]) == 150


def test_c2x2_ends_on_a_window_boundary(datasets):
    ds = datasets("c2x2")
    seq = "".join(l.strip() for l in open(ds.fasta) if not l.startswith(">"))
    assert len(seq) == 2 * 262144


def test_generation_is_independent_of_thread_count(built, tmp_path):
    code = ("import sys, gsnp_amd; gsnp_amd.synth(3, 14, 0.002, sys.argv[1]); gsnp_amd.synth(6, 13, 0.3, sys.argv[2])")
    outs = []
    for t in ("1", "5"):
        env = dict(os.environ, GSNP_SYNTH_THREADS=t)
        a, b = tmp_path / f"c3_{t}", tmp_path / f"st_{t}"
        subprocess.run([sys.executable, "-c", code, str(a), str(b)], check=True, env=env,
                       cwd=Path(__file__).resolve().parents[1])
        outs.append([Path(f"{p}{s}").read_bytes() for p in (a, b) for s in (".fa", ".soap")])
    assert outs[0] == outs[1]
