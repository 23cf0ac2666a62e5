"""The -C reader (gsnp_decompress, gsnp_cns_open/next) on damaged input and
the writers on a full disk (CPU only).

Valid -C files come from the oracle on a small seeded data set.  Truncated
files, flipped bits and hand-made hostile headers must come back as
GSNP_E_FORMAT, never as a crash, an out-of-range write or a huge allocation;
each damaged decode runs in a child process (the gsnp-decompress tool) so a
crash would show as a signal, not take the test runner down.  Writing to
/dev/full must give GSNP_E_IO."""
from __future__ import annotations

import ctypes
import os
import shutil
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_lib

SUFFIXES = [".1_2", ".3_4_6", ".5_7_8_9_14_16", ".10", ".11_12_13", ".15_17"]
BINARY = SUFFIXES[1:]
E_IO, E_FORMAT = 1, 2


@pytest.fixture(scope="module")
def cfiles(built, tmp_path_factory):
    """Oracle -C files of a 40 kb single-contig set (one window) and of the
    3-contig stress set (several windows, known SNPs, N runs)."""
    root = tmp_path_factory.mktemp("cfiles")
    out = {}
    for name, cfg, seed, scale in (("small", 1, 21, 0.4), ("stress", 6, 13, 0.3)):
        pre = root / name
        _, rl = built.synth(cfg, seed, scale, pre)
        snp = Path(f"{pre}.snp")
        oracle_lib.run(f"{pre}.fa", f"{pre}.soap", root / f"{name}.c", dbsnp=snp if snp.exists() else None,
                       compress=True, read_length=rl, rank_sum_mode=1)
        oracle_lib.run(f"{pre}.fa", f"{pre}.soap", root / f"{name}.cns", dbsnp=snp if snp.exists() else None,
                       read_length=rl, rank_sum_mode=1)
        out[name] = (root / f"{name}.c", root / f"{name}.cns")
    return out


def _copy(src: Path, dst: Path) -> None:
    for s in SUFFIXES:
        shutil.copyfile(f"{src}{s}", f"{dst}{s}")


def _decompress(built, prefix: Path, out: Path) -> int:
    """Exit status of gsnp-decompress: 0 ok, 1 reported error, negative a signal."""
    r = subprocess.run([str(built.DECOMPRESS_PATH), str(prefix), str(out)], capture_output=True, text=True, timeout=60)
    return r.returncode


def test_intact_files_decompress_to_the_oracle_text(built, cfiles, tmp_path):
    for name, (pre, cns) in cfiles.items():
        out = tmp_path / f"{name}.out"
        assert built.lib().gsnp_decompress(os.fsencode(pre), os.fsencode(out)) == 0, built.lib().gsnp_last_error()
        assert out.read_bytes() == cns.read_bytes(), name


@pytest.mark.parametrize("sfx", BINARY)
def test_truncated_files_are_format_errors(built, cfiles, tmp_path, sfx):
    pre, _ = cfiles["stress"]
    size = os.path.getsize(f"{pre}{sfx}")
    cuts = sorted({0, 1, 4, 7, 8, 55, 56, 57, size // 3, size // 2, size - 9, size - 4, size - 1} & set(range(size)))
    for cut in cuts:
        dst = tmp_path / f"t{cut}"
        _copy(pre, dst)
        with open(f"{dst}{sfx}", "r+b") as f:
            f.truncate(cut)
        rc = built.lib().gsnp_decompress(os.fsencode(dst), os.fsencode(tmp_path / "o.cns"))
        assert rc == E_FORMAT, (sfx, cut, rc, built.lib().gsnp_last_error())


@pytest.mark.parametrize("sfx", BINARY)
def test_flipped_bits_never_crash(built, cfiles, tmp_path, sfx):
    """Random single-bit flips: the decoder either succeeds (most flips land
    in packed values and only change them) or reports a format error, in a
    child process that must not die by a signal."""
    pre, _ = cfiles["small"]
    data = Path(f"{pre}{sfx}").read_bytes()
    rng = np.random.default_rng(1000 + BINARY.index(sfx))
    dst = tmp_path / "flip"
    _copy(pre, dst)
    for trial in range(40):
        b = bytearray(data)
        pos = int(rng.integers(0, len(b)))
        b[pos] ^= 1 << int(rng.integers(0, 8))
        Path(f"{dst}{sfx}").write_bytes(bytes(b))
        rc = _decompress(built, dst, tmp_path / "o.cns")
        assert rc in (0, 1), (sfx, trial, pos, rc)


def _first_header(built, cfiles, tmp_path, sfx, name="small"):
    pre, _ = cfiles[name]
    dst = tmp_path / "h"
    _copy(pre, dst)
    return dst, bytearray(Path(f"{dst}{sfx}").read_bytes())


@pytest.mark.parametrize("case", ["n4_huge", "n6_past_window", "id4_outside", "dict_index_past_dict", "rle_runs_short",
                                  "rle_nword_wrong", "col10_index_outside", "col11_common_huge", "col15_negative",
                                  "col15_index_outside", "window_size_huge"])
def test_hostile_headers_are_format_errors(built, cfiles, tmp_path, case):
    """Counts and indices edited to point outside the window or the
    dictionaries: each must be refused before it is used."""
    L = built.lib()
    if case in ("n4_huge", "n6_past_window", "id4_outside"):
        dst, b = _first_header(built, cfiles, tmp_path, ".3_4_6")
        pa = list(struct.unpack_from("<7Q", b, 0))
        assert pa[0] == 0 and pa[5] > 0, "the first window should have col-4 differences"
        if case == "n4_huge":
            struct.pack_into("<Q", b, 5 * 8, (1 << 63) + 5)
        elif case == "n6_past_window":
            struct.pack_into("<Q", b, 6 * 8, pa[2] + 1)
        else:
            struct.pack_into("<i", b, 56, pa[2] + 3)  # first col-4 id
        Path(f"{dst}.3_4_6").write_bytes(bytes(b))
    elif case in ("dict_index_past_dict", "rle_runs_short", "rle_nword_wrong"):
        dst, b = _first_header(built, cfiles, tmp_path, ".5_7_8_9_14_16")
        nind, = struct.unpack_from("<i", b, 0)
        assert nind == 0
        pa = list(struct.unpack_from("<6Q", b, 4))
        if case == "dict_index_past_dict":
            # one dictionary entry fewer: the largest index no longer has an entry
            # (the header stays consistent in size by keeping the word count)
            nd = pa[3]
            assert nd >= 3
            struct.pack_into("<Q", b, 4 + 3 * 8, nd - 1)
            struct.pack_into("<Q", b, 4 + 5 * 8, pa[5] - 4)
            # drop the last dictionary entry's 4 bytes from the blob
            blob_start = 4 + 48
            cut = blob_start + 4 * pa[1] + 4 * pa[2] + 4 * (nd - 1)
            b = b[:cut] + b[cut + 4:]
        elif case == "rle_runs_short":
            # first run-length dictionary value shrunk: the runs no longer cover the window
            off = 4 + 48 + 4 * pa[1] + 4 * pa[2] + 4 * pa[3]
            v, = struct.unpack_from("<I", b, off)
            struct.pack_into("<I", b, off, max(0, v - 1) if v > 1 else v + 10 ** 6)
        else:
            struct.pack_into("<Q", b, 4 + 1 * 8, pa[1] + 1)
        Path(f"{dst}.5_7_8_9_14_16").write_bytes(bytes(b))
    elif case == "col10_index_outside":
        dst, b = _first_header(built, cfiles, tmp_path, ".10", "stress")
        p3 = struct.unpack_from("<3Q", b, 4)
        assert p3[0] > 0
        struct.pack_into("<I", b, 4 + 24 + 4 * p3[1], p3[2] + 7)
        Path(f"{dst}.10").write_bytes(bytes(b))
    elif case == "col11_common_huge":
        dst, b = _first_header(built, cfiles, tmp_path, ".11_12_13")
        struct.pack_into("<Q", b, 4 + 8, 1 << 62)
        Path(f"{dst}.11_12_13").write_bytes(bytes(b))
    elif case in ("col15_negative", "col15_index_outside"):
        dst, b = _first_header(built, cfiles, tmp_path, ".15_17")
        n, = struct.unpack_from("<i", b, 4)
        if case == "col15_negative":
            struct.pack_into("<i", b, 4, -5)
        else:
            assert n > 0
            struct.pack_into("<i", b, 8, 1 << 30)
        Path(f"{dst}.15_17").write_bytes(bytes(b))
    else:
        pre, _ = cfiles["small"]
        dst = tmp_path / "h"
        _copy(pre, dst)
        lines = Path(f"{dst}.1_2").read_text().splitlines()
        name, start, size = lines[1].split("\t")
        lines[1] = f"{name}\t{start}\t{1 << 40}"
        Path(f"{dst}.1_2").write_text("\n".join(lines) + "\n")
    rc = L.gsnp_decompress(os.fsencode(dst), os.fsencode(tmp_path / "o.cns"))
    assert rc == E_FORMAT, (case, rc, L.gsnp_last_error())
    # the in-memory reader refuses it the same way
    r = ctypes.c_void_p()
    assert L.gsnp_cns_open(os.fsencode(dst), ctypes.byref(r)) == 0
    try:
        w = built.CnsWindow()
        has = ctypes.c_int(1)
        rc = 0
        while rc == 0 and has.value:
            rc = L.gsnp_cns_next(r, ctypes.byref(w), ctypes.byref(has))
        assert rc == E_FORMAT, case
    finally:
        L.gsnp_cns_close(r)


def test_decompressing_to_a_full_disk_is_an_io_error(built, cfiles):
    if not os.path.exists("/dev/full"):
        pytest.skip("no /dev/full")
    pre, _ = cfiles["stress"]
    rc = built.lib().gsnp_decompress(os.fsencode(pre), b"/dev/full")
    assert rc == E_IO, built.lib().gsnp_last_error()
