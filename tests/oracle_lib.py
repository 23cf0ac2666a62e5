"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) and of the
reference's own dictionary codec built from cpp/dict_compression.h
(oracle/_ref/libdictref.so).  Test infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "_build" / "liboracle.so"
ORACLE_CLI = ORACLE_DIR / "_build" / "gsnp_oracle"
DICTREF_SO = ORACLE_DIR / "_ref" / "libdictref.so"


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("q_min", ctypes.c_int), ("q_max", ctypes.c_int), ("read_length", ctypes.c_int),
        ("is_monoploid", ctypes.c_int), ("refine_mode", ctypes.c_int), ("rank_sum_mode", ctypes.c_int),
        ("transition_dominant", ctypes.c_int),
        ("althom_novel_r", ctypes.c_double), ("het_novel_r", ctypes.c_double),
        ("althom_val_r", ctypes.c_double), ("het_val_r", ctypes.c_double),
        ("althom_unval_r", ctypes.c_double), ("het_unval_r", ctypes.c_double),
        ("global_dependency", ctypes.c_double), ("pcr_dependency", ctypes.c_double),
    ]


class OrcSnp(ctypes.Structure):
    _fields_ = [("hapmap", ctypes.c_uint8), ("validated", ctypes.c_uint8), ("indel", ctypes.c_uint8),
                ("freq", ctypes.c_double * 4)]


class OrcSite(ctypes.Structure):
    _fields_ = [
        ("depth", ctypes.c_int32), ("repeat", ctypes.c_int32), ("dep_uni", ctypes.c_int32),
        ("cnt_uni", ctypes.c_int32 * 4), ("q_sum", ctypes.c_int32 * 4), ("cnt_all", ctypes.c_int32 * 4),
        ("ori", ctypes.c_uint8), ("nflag", ctypes.c_uint8), ("base1", ctypes.c_uint8), ("base2", ctypes.c_uint8),
        ("type1", ctypes.c_uint8), ("type2", ctypes.c_uint8),
        ("qual1", ctypes.c_int32), ("qual2", ctypes.c_int32), ("qual3", ctypes.c_int32), ("q_cns", ctypes.c_int32),
        ("rank_sum", ctypes.c_double), ("lk", ctypes.c_double * 10),
    ]


class OrcReads(ctypes.Structure):
    """orc_reads: the parsed alignments (seq/qual hold L chars per read)."""

    _fields_ = [
        ("n", ctypes.c_int64), ("L", ctypes.c_int),
        ("pos", ctypes.POINTER(ctypes.c_int32)), ("len", ctypes.POINTER(ctypes.c_int32)),
        ("hit", ctypes.POINTER(ctypes.c_int32)), ("strand", ctypes.POINTER(ctypes.c_uint8)),
        ("chr", ctypes.POINTER(ctypes.c_int32)),
        ("seq", ctypes.POINTER(ctypes.c_char)), ("qual", ctypes.POINTER(ctypes.c_char)),
    ]


_orc = None
_ref = None


def ensure_built() -> None:
    if not ORACLE_SO.exists() or not ORACLE_CLI.exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
    if not DICTREF_SO.exists() and Path("/root/reference/cpp/dict_compression.h").exists():
        subprocess.run(["make", "-C", str(ORACLE_DIR), "ref"], check=True, capture_output=True)


def orc() -> ctypes.CDLL:
    global _orc
    if _orc is None:
        ensure_built()
        L = ctypes.CDLL(str(ORACLE_SO))
        d, i, u32, u64, P = ctypes.c_double, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p
        sig = {
            "orc_params_default": (None, [ctypes.POINTER(OrcParams)]),
            "orc_encode_word": (u32, [u32, u32, u32, u32, u32]),
            "orc_sort_records": (None, [P, i]),
            "orc_likelihood_sorted": (i, [P, i, i, d, d, P, P]),
            "orc_likelihood_dense": (i, [P, i, i, i, d, d, P, P]),
            "orc_qadj": (i, [i, i, i, d, d]),
            "orc_normal_value": (d, [d]),
            "orc_normal_test": (d, [i, i, d, d]),
            "orc_table_test": (d, [P, i, i, d, d]),
            "orc_rank_test": (d, [P, i, P, i, i, P]),
            "orc_prior_gen": (None, [ctypes.POINTER(OrcParams), P]),
            "orc_rank_table_gen": (None, [P]),
            "orc_score_table": (None, [P, P]),
            "orc_format_g": (i, [d, ctypes.c_char_p]),
            "orc_trunc_int": (i, [d]),
            "orc_num_bits": (i, [u64]),
            "orc_dict_encode_i32": (u64, [P, u64, P, P, ctypes.POINTER(u64)]),
            "orc_dict_encode_f64": (u64, [P, u64, P, P, ctypes.POINTER(u64)]),
            "orc_encode_2bit": (u64, [ctypes.c_char_p, u64, P]),
            "orc_decode_2bit": (None, [ctypes.c_char_p, u64, P, u64]),
            "orc_posterior_site": (None, [ctypes.POINTER(OrcSite), P, i, P, ctypes.POINTER(OrcParams), P,
                                          ctypes.POINTER(OrcSnp), P]),
            "orc_snp_prior": (i, [P, i, ctypes.POINTER(OrcSnp), ctypes.POINTER(OrcParams)]),
            "orc_run": (i, [ctypes.c_char_p] * 5 + [ctypes.POINTER(OrcParams), ctypes.c_char_p, i, ctypes.c_int64, P]),
            "orc_decompress": (i, [ctypes.c_char_p, ctypes.c_char_p]),
            "orc_last_error": (ctypes.c_char_p, []),
            "orc_genome_load": (i, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(P)]),
            "orc_genome_free": (None, [P]),
            "orc_reads_load": (i, [ctypes.c_char_p, ctypes.POINTER(OrcParams), P, ctypes.POINTER(P)]),
            "orc_reads_free": (None, [P]),
            "orc_matrix_gen": (i, [P, P, ctypes.POINTER(OrcParams), P]),
            "orc_format_row": (i, [ctypes.c_char_p, u32, ctypes.POINTER(OrcSite), ctypes.POINTER(OrcSite),
                                   ctypes.c_char_p]),
            "orc_cwrite_sites": (i, [ctypes.c_char_p, ctypes.c_char_p, u32, u32, ctypes.POINTER(OrcSite)]),
            "orc_region_set": (i, [P, P, u32, i, i]),
            "orc_region_in_win": (i, [P, u32]),
            "orc_region_words": (u64, [u64]),
            "orc_run_opts": (i, [ctypes.c_char_p] * 5 + [ctypes.POINTER(OrcParams), ctypes.c_char_p, i,
                                 ctypes.c_int64, i, ctypes.c_char_p, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _orc = L
    return _orc


def dictref() -> ctypes.CDLL | None:
    """The reference's dictionary codec, or None where it cannot be built."""
    global _ref
    if _ref is None:
        ensure_built()
        if not DICTREF_SO.exists():
            return None
        L = ctypes.CDLL(str(DICTREF_SO))
        P, u64 = ctypes.c_void_p, ctypes.c_ulong
        L.ref_num_bits.restype = ctypes.c_int
        L.ref_num_bits.argtypes = [u64]
        L.ref_dict_encode_i32.restype = u64
        L.ref_dict_encode_i32.argtypes = [P, u64, P, P, ctypes.POINTER(u64)]
        L.ref_dict_encode_f64.restype = u64
        L.ref_dict_encode_f64.argtypes = [P, u64, P, P, ctypes.POINTER(u64)]
        L.ref_dict_encode_char.restype = u64
        L.ref_dict_encode_char.argtypes = [P, u64, P, P, ctypes.POINTER(u64)]
        L.ref_dict_encode_u32.restype = u64
        L.ref_dict_encode_u32.argtypes = [P, u64, P, P, ctypes.POINTER(u64)]
        L.ref_dict_decode_i32.restype = None
        L.ref_dict_decode_i32.argtypes = [P, u64, P, u64, P, u64]
        _ref = L
    return _ref


def params(**kw) -> OrcParams:
    import math
    p = OrcParams()
    orc().orc_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("global_dependency", "pcr_dependency"):
            v = math.log10(v)
        setattr(p, k, v)
    return p


def run(fasta, soap, out, dbsnp=None, compress=False, max_sites=0, regions=None, **kw) -> None:
    p = params(**kw)
    b = lambda s: None if s is None else os.fsencode(s)
    rc = orc().orc_run_opts(b(fasta), b(soap), b(dbsnp), None, None, ctypes.byref(p), b(out), int(compress),
                            max_sites, 1, b(regions), None)
    if rc:
        raise RuntimeError(f"oracle error {rc}: {orc().orc_last_error().decode()}")


def matrix_gen(fasta, soap, dbsnp=None, **kw):
    """The oracle's counting-trained calibration matrix (256*256*16 doubles)."""
    import numpy as np

    L = orc()
    p = params(**kw)
    b = lambda s: None if s is None else os.fsencode(s)
    g, r = ctypes.c_void_p(), ctypes.c_void_p()
    rc = L.orc_genome_load(b(fasta), b(dbsnp), ctypes.byref(g))
    if rc:
        raise RuntimeError(f"oracle error {rc}: {L.orc_last_error().decode()}")
    try:
        rc = L.orc_reads_load(b(soap), ctypes.byref(p), g, ctypes.byref(r))
        if rc:
            raise RuntimeError(f"oracle error {rc}: {L.orc_last_error().decode()}")
        pm = np.empty(256 * 256 * 16, dtype=np.float64)
        rc = L.orc_matrix_gen(r, g, ctypes.byref(p), pm.ctypes.data)
        L.orc_reads_free(r)
        if rc:
            raise RuntimeError(f"oracle error {rc}: {L.orc_last_error().decode()}")
        return pm
    finally:
        L.orc_genome_free(g)


def decompress(prefix, out) -> None:
    rc = orc().orc_decompress(os.fsencode(prefix), os.fsencode(out))
    if rc:
        raise RuntimeError(f"oracle error {rc}: {orc().orc_last_error().decode()}")
