"""The command-line drop-in: the gsnp and gsnp-decompress binaries against the
oracle's command line on the same flags (cpp/main.cc:104-316).

Each case runs both tools as child processes and compares their output files
byte for byte: -u, -2 -s, -t, -g/-p (given as values, stored log10), -z (with
q_max = q_min + 40) and -Q, -L, -m and the prior flags, -M/-I round trips,
-C with a decompression round trip, the batched path (-B) and -T target
regions (which the reference refuses).  The flags the reference refuses and
the gsnp binary refuses too (-q, -F) exit 0 with its message and write nothing."""
from __future__ import annotations

import filecmp
import subprocess
from pathlib import Path

import pytest

import oracle_lib

pytestmark = pytest.mark.gpu

SUFFIXES = [".1_2", ".3_4_6", ".5_7_8_9_14_16", ".10", ".11_12_13", ".15_17"]


def _gsnp(built, args, timeout=300):
    print(f"[cli] gsnp {' '.join(map(str, args))}", flush=True)
    return subprocess.run([str(built.CLI_PATH)] + [str(a) for a in args], capture_output=True, text=True,
                          timeout=timeout)


def _oracle(args, timeout=600):
    oracle_lib.ensure_built()
    return subprocess.run([str(oracle_lib.ORACLE_CLI)] + [str(a) for a in args], capture_output=True, text=True,
                          timeout=timeout)


def _inputs(ds):
    a = ["-i", ds.soap, "-d", ds.fasta, "-L", ds.read_len]
    if ds.dbsnp:
        a += ["-s", ds.dbsnp]
    return a


CASES = {
    "rank_refine_ti": ["-u", "-2", "-t"],
    "dependencies": ["-u", "-g", "0.75", "-p", "0.3"],
    "no_dependency": ["-g", "1", "-p", "1"],
    "qmin_shift": ["-z", "<", "-2"],  # q_min 60: the stress set has quality characters down to 61
    "qmax": ["-Q", "d", "-u"],
    "priors_monoploid": ["-m", "-r", "0.001", "-e", "0.002", "-a", "0.06", "-b", "0.11", "-j", "0.02", "-k", "0.03",
                         "-2"],
}


@pytest.mark.parametrize("case", list(CASES))
def test_cli_text_matches_oracle_cli(built, datasets, tmp_path, case):
    ds = datasets("stress")
    flags = CASES[case]
    g, o = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    r = _gsnp(built, _inputs(ds) + flags + ["-o", g])
    assert r.returncode == 0, r.stderr
    r2 = _oracle(_inputs(ds) + flags + ["-o", o])
    assert r2.returncode == 0, r2.stderr
    assert g.read_bytes() == o.read_bytes(), case


def test_cli_qmin_flag_changes_the_calls(built, datasets, tmp_path):
    """-z shifts every quality by four here, so the text must differ from the
    default run (a flag that were ignored would pass the comparison above)."""
    ds = datasets("stress")
    a, b = tmp_path / "a.cns", tmp_path / "b.cns"
    assert _gsnp(built, _inputs(ds) + ["-o", a]).returncode == 0
    assert _gsnp(built, _inputs(ds) + ["-z", "<", "-o", b]).returncode == 0
    assert a.read_bytes() != b.read_bytes()


def test_cli_matrix_write_then_read(built, datasets, tmp_path):
    """-M writes the oracle's matrix text; -I of that file calls as -M did."""
    ds = datasets("stress")
    m_gpu, m_cpu = tmp_path / "gpu.matrix", tmp_path / "cpu.matrix"
    c1, c2, c3 = tmp_path / "m.cns", tmp_path / "i.cns", tmp_path / "cpu.cns"
    r = _gsnp(built, _inputs(ds) + ["-u", "-M", m_gpu, "-o", c1])
    assert r.returncode == 0, r.stderr
    assert _oracle(_inputs(ds) + ["-u", "-M", m_cpu, "-o", c3]).returncode == 0
    assert m_gpu.read_bytes() == m_cpu.read_bytes()
    lines = m_gpu.read_text().splitlines()
    assert len(lines) == 41 * ds.read_len  # q 0..40 (q_max = q_min + 40) by cycle 0..L-1
    assert len(lines[0].split("\t")) == 18
    r = _gsnp(built, _inputs(ds) + ["-u", "-I", m_gpu, "-o", c2])
    assert r.returncode == 0, r.stderr
    assert c1.read_bytes() == c2.read_bytes() == c3.read_bytes()


def test_cli_matrix_flags_share_one_stream(built, datasets, tmp_path):
    """-M then -I: the matrix is read from -I and the -M file is left empty;
    -I then -M is refused (SOAPsnp would read no matrix at all)."""
    ds = datasets("stress")
    m = tmp_path / "x.matrix"
    assert _gsnp(built, _inputs(ds) + ["-M", m, "-o", tmp_path / "a.cns"]).returncode == 0
    y = tmp_path / "y.matrix"
    r = _gsnp(built, _inputs(ds) + ["-M", y, "-I", m, "-o", tmp_path / "b.cns"])
    assert r.returncode == 0, r.stderr
    assert y.exists() and y.stat().st_size == 0
    assert (tmp_path / "a.cns").read_bytes() == (tmp_path / "b.cns").read_bytes()
    r = _gsnp(built, _inputs(ds) + ["-I", m, "-M", y, "-o", tmp_path / "c.cns"])
    assert r.returncode == 255


@pytest.mark.parametrize("name,extra", [("stress", ["-u", "-2"]), ("c5s", ["-u"])])
def test_cli_compressed_and_decompressed(built, datasets, tmp_path, name, extra):
    ds = datasets(name)
    g, o = tmp_path / "gpu", tmp_path / "cpu"
    r = _gsnp(built, _inputs(ds) + extra + ["-C", "-o", g])
    assert r.returncode == 0, r.stderr
    assert _oracle(_inputs(ds) + extra + ["-C", "-o", o]).returncode == 0
    for s in SUFFIXES:
        assert filecmp.cmp(f"{g}{s}", f"{o}{s}", shallow=False), s
    txt = tmp_path / "dec.cns"
    r = subprocess.run([str(built.DECOMPRESS_PATH), str(g), str(txt)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    plain = tmp_path / "plain.cns"
    assert _oracle(_inputs(ds) + extra + ["-o", plain]).returncode == 0
    assert txt.read_bytes() == plain.read_bytes()


def test_cli_batched_matches_plain(built, datasets, tmp_path):
    ds = datasets("c3s")
    a, b = tmp_path / "plain.cns", tmp_path / "batched.cns"
    assert _gsnp(built, _inputs(ds) + ["-u", "-o", a]).returncode == 0
    assert _gsnp(built, _inputs(ds) + ["-u", "-B", "8", "-o", b]).returncode == 0
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("flag", [["-q"], ["-F", "1"]])
def test_cli_refused_flags_exit_quietly(built, datasets, tmp_path, flag):
    ds = datasets("c1")
    out = tmp_path / "x.cns"
    r = _gsnp(built, _inputs(ds) + flag + ["-o", out])
    assert r.returncode == 0
    assert "currently is not supported" in r.stdout
    assert not out.exists()


def test_cli_missing_compulsory_arguments(built, tmp_path):
    r = _gsnp(built, ["-o", tmp_path / "x.cns"])
    assert r.returncode != 0


def test_cli_write_failure_is_reported(built, datasets):
    ds = datasets("c1")
    r = _gsnp(built, _inputs(ds) + ["-o", "/dev/full"])
    assert r.returncode == 255 and "cannot write" in r.stderr


def test_cli_target_regions(built, datasets, tmp_path):
    """-T (refused by the reference) calls the windows the regions touch, as the oracle does."""
    ds = datasets("stress")
    reg = tmp_path / "r.txt"
    reg.write_text("chr1\t300000\t300050\nchr3\t10\t20\n")
    g, o = tmp_path / "gpu.cns", tmp_path / "cpu.cns"
    r = _gsnp(built, _inputs(ds) + ["-u", "-T", reg, "-o", g])
    assert r.returncode == 0, r.stderr
    assert _oracle(_inputs(ds) + ["-u", "-T", reg, "-o", o]).returncode == 0
    assert g.read_bytes() == o.read_bytes() and g.stat().st_size > 0
