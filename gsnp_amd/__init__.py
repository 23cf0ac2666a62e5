"""gsnp_amd -- MI355X (gfx950) SOAPsnp SNP-calling path.

The product is the C-ABI library ``gsnp_amd/_build/libgsnp.so`` (declared in
``include/gsnp_abi.h``) with HIP kernels for the site matrix, segmented sort,
FP64 genotype likelihood, posterior, rank-sum test and .cns formatting, plus
the command-line tools ``gsnp`` and ``gsnp-decompress`` beside it.  This
module is a thin ctypes binding for Python callers; it loads the built
library and raises if it is missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import math
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "_build" / "libgsnp.so"
CLI_PATH = PKG / "_build" / "gsnp"
DECOMPRESS_PATH = PKG / "_build" / "gsnp-decompress"
WIN = 262144

GSNP_OK = 0
ERRORS = {1: "io", 2: "format", 3: "unsorted", 4: "chrom", 5: "qual", 6: "range", 7: "nomem", 8: "arg", 9: "device"}


class GsnpError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gsnp error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class Params(ctypes.Structure):
    """Mirror of gsnp_params (cpp/common.h:123-154)."""

    _fields_ = [
        ("q_min", ctypes.c_int),
        ("q_max", ctypes.c_int),
        ("read_length", ctypes.c_int),
        ("is_monoploid", ctypes.c_int),
        ("refine_mode", ctypes.c_int),
        ("rank_sum_mode", ctypes.c_int),
        ("transition_dominant", ctypes.c_int),
        ("althom_novel_r", ctypes.c_double),
        ("het_novel_r", ctypes.c_double),
        ("althom_val_r", ctypes.c_double),
        ("het_val_r", ctypes.c_double),
        ("althom_unval_r", ctypes.c_double),
        ("het_unval_r", ctypes.c_double),
        ("global_dependency", ctypes.c_double),
        ("pcr_dependency", ctypes.c_double),
    ]


class WindowOut(ctypes.Structure):
    _fields_ = [
        ("base1", ctypes.c_void_p),
        ("base2", ctypes.c_void_p),
        ("type1", ctypes.c_void_p),
        ("q_cns", ctypes.c_void_p),
        ("rank_sum", ctypes.c_void_p),
        ("lk10", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("dep_uni", ctypes.c_void_p),
    ]


class ReadBatch(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_uint64),
        ("pitch", ctypes.c_uint32),
        ("pos", ctypes.c_void_p),
        ("len_hit", ctypes.c_void_p),
        ("strand", ctypes.c_void_p),
        ("seq", ctypes.c_void_p),
        ("qual", ctypes.c_void_p),
    ]


class Snp(ctypes.Structure):
    _fields_ = [
        ("hapmap", ctypes.c_uint8),
        ("validated", ctypes.c_uint8),
        ("indel", ctypes.c_uint8),
        ("pad", ctypes.c_uint8),
        ("freq", ctypes.c_double * 4),
    ]


_lib = None

# every entry point declared in include/gsnp_abi.h: (restype, argtypes)
_P = ctypes.c_void_p
SIGNATURES = {
    "gsnp_params_default": (None, [ctypes.POINTER(Params)]),
    "gsnp_last_error": (ctypes.c_char_p, []),
    "gsnp_device_init": (ctypes.c_int, [ctypes.c_int]),
    "gsnp_create": (ctypes.c_int, [ctypes.POINTER(Params), _P, ctypes.POINTER(_P)]),
    "gsnp_destroy": (None, [_P]),
    "gsnp_set_chromosome": (ctypes.c_int, [_P, ctypes.c_char_p, _P, ctypes.c_uint32, _P, _P, ctypes.c_uint32]),
    "gsnp_load_reads": (ctypes.c_int, [_P, ctypes.POINTER(ReadBatch)]),
    "gsnp_window_init": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gsnp_window_update": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32)]),
    "gsnp_window_likelihood": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gsnp_window_posterior": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "gsnp_window_text": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_fetch_text": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_window_results": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.POINTER(WindowOut)]),
    "gsnp_window_recycle": (ctypes.c_int, [_P]),
    "gsnp_window_columns": (ctypes.c_int, [_P, ctypes.c_uint32, _P, _P, _P]),
    "gsnp_set_stream": (ctypes.c_int, [_P, _P]),
    "gsnp_job_columns_to": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_batched_call": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_run": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                ctypes.POINTER(Params), ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    "gsnp_decompress": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p]),
    "gsnp_decompress_device": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p]),
    "gsnp_matrix_gen": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(Params), _P]),
    "gsnp_dict_compress_i32": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, ctypes.POINTER(ctypes.c_uint64), _P,
                                              ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_dict_compress_char": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, ctypes.POINTER(ctypes.c_uint64), _P,
                                               ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_dict_compress_u32": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, ctypes.POINTER(ctypes.c_uint64), _P,
                                              ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_dict_compress_f64": (ctypes.c_int, [_P, _P, ctypes.c_uint64, _P, ctypes.POINTER(ctypes.c_uint64), _P,
                                              ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_sort_likelihood": (ctypes.c_int, [_P, _P, ctypes.c_uint32, _P, ctypes.c_uint32, _P]),
    "gsnp_sort_likelihood_noalloc": (ctypes.c_int, [_P, _P, _P, ctypes.c_uint32, _P]),
    "gsnp_format_g": (ctypes.c_int, [ctypes.c_double, ctypes.c_char_p]),
    "gsnp_profile": (ctypes.c_int, [_P, ctypes.c_int, _P, _P, ctypes.c_int]),
    "gsnp_stats": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.c_int]),
    "gsnp_counters": (ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int]),
    "gsnp_job_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                     ctypes.POINTER(Params), ctypes.POINTER(_P)]),
    "gsnp_job_count": (ctypes.c_int, [_P]),
    "gsnp_job_info": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_char_p, ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_job_select": (ctypes.c_int, [_P, ctypes.c_int]),
    "gsnp_job_call": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_job_columns": (ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_job_write": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int, ctypes.c_int]),
    "gsnp_job_set_regions": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gsnp_job_write_matrix": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "gsnp_job_pass": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint64)]),
    "gsnp_job_synth": (ctypes.c_int, [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, ctypes.POINTER(Params),
                                      ctypes.c_char_p, ctypes.POINTER(_P), ctypes.POINTER(ctypes.c_int)]),
    "gsnp_set_matrix": (ctypes.c_int, [_P, _P]),
    "gsnp_cns_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(_P)]),
    "gsnp_cns_next": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_int)]),
    "gsnp_cns_close": (None, [_P]),
    "gsnp_calib_count": (ctypes.c_int, [_P, _P]),
    "gsnp_job_context": (_P, [_P]),
    "gsnp_job_close": (None, [_P]),
    "gsnp_synth_write": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, ctypes.c_char_p,
                                             ctypes.POINTER(ctypes.c_int)]),
    "gsnp_synth_write_contigs": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double,
                                                     ctypes.c_char_p, _P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gsnp_synth_contigs": (ctypes.c_int, [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_double, _P, ctypes.c_int]),
    "gsnp_autotune": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, _P]),
    "gsnp_adjoint_prepare": (ctypes.c_int, [_P, ctypes.c_uint32, ctypes.c_uint32]),
    "gsnp_adjoint_eval": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.POINTER(ctypes.c_double), _P,
                                         ctypes.POINTER(ctypes.c_double), _P]),
    "gsnp_adjoint_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64),
                                         ctypes.POINTER(ctypes.c_uint32)]),
    "gsnp_adjoint_export": (ctypes.c_int, [_P, _P, _P, _P]),
    "gsnp_adjoint_priors": (ctypes.c_int, [_P, ctypes.c_double, _P, _P, ctypes.c_uint32]),
}


def lib() -> ctypes.CDLL:
    """The loaded libgsnp.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != GSNP_OK:
        raise GsnpError(rc, lib().gsnp_last_error().decode())


def params(**kw) -> Params:
    """Default SOAPsnp parameters (cpp/common.h:142-153), overridden by kw.
    global_dependency / pcr_dependency are given as the -g / -p values and
    stored log10'd, as cpp/main.cc:151-160 does."""
    p = Params()
    lib().gsnp_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k in ("global_dependency", "pcr_dependency"):
            v = math.log10(v)
        setattr(p, k, v)
    return p


def _b(s) -> bytes | None:
    return None if s is None else os.fsencode(s)


def run(fasta, soap, out, dbsnp=None, matrix_in=None, matrix_out=None, compress=False, batched=0, **kw) -> None:
    """Whole pipeline (gsnp_run): FASTA + SOAP (+ known SNPs) to .cns / -C files."""
    L = lib()
    check(L.gsnp_device_init(0))
    p = params(**kw)
    check(L.gsnp_run(_b(fasta), _b(soap), _b(dbsnp), _b(matrix_in), _b(matrix_out), ctypes.byref(p), _b(out),
                     int(bool(compress)), int(batched)))


def decompress(prefix, out, device: bool = False) -> None:
    """-C files back to .cns text, on the host or (device=True) on the GPU."""
    L = lib()
    if device:
        check(L.gsnp_device_init(0))
        check(L.gsnp_decompress_device(_b(prefix), _b(out)))
    else:
        check(L.gsnp_decompress(_b(prefix), _b(out)))


class CnsWindow(ctypes.Structure):
    """One window of -C output as its 17 .cns columns (gsnp_cns_window)."""

    _c, _i, _d = ctypes.POINTER(ctypes.c_char), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)
    _fields_ = [("chr", ctypes.c_char_p), ("start", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("ref", _c), ("cns", _c), ("q_cns", _i), ("base1", _c), ("avg_q1", _i), ("count_uni1", _i),
                ("count_all1", _i), ("base2", _c), ("avg_q2", _i), ("count_uni2", _i), ("count_all2", _i),
                ("depth", _i), ("rank_sum", _d), ("copy_num", _d), ("dbsnp", _i)]


def cns_windows(prefix):
    """The windows of -C files, read in memory one at a time (gsnp_cns_open /
    gsnp_cns_next): yields (chr, start, columns), the columns a dict of numpy
    arrays named after gsnp_cns_window's fields (copies)."""
    import numpy as np

    L = lib()
    r = ctypes.c_void_p()
    check(L.gsnp_cns_open(_b(prefix), ctypes.byref(r)))
    try:
        w = CnsWindow()
        has = ctypes.c_int()
        while True:
            check(L.gsnp_cns_next(r, ctypes.byref(w), ctypes.byref(has)))
            if not has.value:
                return
            n = int(w.n)
            cols = {}
            for name, ctype, _ in (("ref", ctypes.c_char, 0), ("cns", ctypes.c_char, 0), ("q_cns", ctypes.c_int32, 0),
                                   ("base1", ctypes.c_char, 0), ("avg_q1", ctypes.c_int32, 0),
                                   ("count_uni1", ctypes.c_int32, 0), ("count_all1", ctypes.c_int32, 0),
                                   ("base2", ctypes.c_char, 0), ("avg_q2", ctypes.c_int32, 0),
                                   ("count_uni2", ctypes.c_int32, 0), ("count_all2", ctypes.c_int32, 0),
                                   ("depth", ctypes.c_int32, 0), ("rank_sum", ctypes.c_double, 0),
                                   ("copy_num", ctypes.c_double, 0), ("dbsnp", ctypes.c_int32, 0)):
                ptr = getattr(w, name)
                dt = {ctypes.c_char: np.uint8, ctypes.c_int32: np.int32, ctypes.c_double: np.float64}[ctype]
                cols[name] = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype * n)).contents).view(dt).copy() \
                    if n else np.zeros(0, dt)
            yield w.chr.decode(), int(w.start), cols
    finally:
        L.gsnp_cns_close(r)


def synth(config: int, seed: int, scale: float, prefix) -> tuple[int, int]:
    """Seeded synthetic inputs for a benchmark configuration; returns (alignment lines, read length)."""
    rl = ctypes.c_int()
    n = lib().gsnp_synth_write(config, seed, scale, _b(prefix), ctypes.byref(rl))
    if n < 0:
        raise GsnpError(1, f"cannot write {prefix}.*")
    return int(n), rl.value


def synth_contigs(config: int, seed: int, scale: float) -> list[int]:
    """Contig lengths of a configuration's data set."""
    n = lib().gsnp_synth_contigs(config, seed, scale, None, 0)
    arr = (ctypes.c_uint64 * max(1, n))()
    lib().gsnp_synth_contigs(config, seed, scale, arr, n)
    return [int(arr[i]) for i in range(n)]


def synth_subset(config: int, seed: int, scale: float, contigs, prefix) -> tuple[int, int]:
    """The files of the listed contigs only; returns (alignment lines, read length)."""
    rl = ctypes.c_int()
    idx = (ctypes.c_int * max(1, len(contigs)))(*contigs)
    n = lib().gsnp_synth_write_contigs(config, seed, scale, _b(prefix), idx, len(contigs), ctypes.byref(rl))
    if n < 0:
        raise GsnpError(1, f"cannot write {prefix}.*")
    return int(n), rl.value


def matrix_gen(fasta, soap, dbsnp=None, **kw):
    """Counting-trained calibration matrix (gsnp_matrix_gen), 256*256*16 doubles
    indexed q<<12 | cycle<<4 | ref<<2 | obs, as a numpy array."""
    import numpy as np

    pm = np.zeros(256 * 256 * 16, dtype=np.float64)
    p = params(**kw)
    check(lib().gsnp_matrix_gen(_b(fasta), _b(soap), _b(dbsnp), ctypes.byref(p), pm.ctypes.data))
    return pm


def format_g(v: float) -> str:
    buf = ctypes.create_string_buffer(64)
    n = lib().gsnp_format_g(v, buf)
    return buf.raw[:n].decode()


STAGES = ["site_build", "sort", "likelihood", "posterior", "rank", "format", "recycle", "fused", "adjoint", "columns",
          "calib", "bucket"]


# gsnp_chunk_fn: (user, chromosome, 1-based start, sites, chunk pointers, chunk bytes) -> 0 to go on
CHUNK_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32,
                            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_uint64))


def window_columns(ctx, call_length: int) -> list[bytes]:
    """The five -C chunks of the window just called on `ctx`
    (gsnp_window_columns): sizes first, then the bytes."""
    L = lib()
    sizes = (ctypes.c_uint64 * 5)()
    check(L.gsnp_window_columns(ctx, int(call_length), None, None, sizes))
    bufs = [ctypes.create_string_buffer(max(1, sizes[f])) for f in range(5)]
    dst = (ctypes.c_void_p * 5)(*[ctypes.addressof(b) for b in bufs])
    cap = (ctypes.c_uint64 * 5)(*[len(b) for b in bufs])
    check(L.gsnp_window_columns(ctx, int(call_length), dst, cap, sizes))
    return [bufs[f].raw[:sizes[f]] for f in range(5)]


class Job:
    """A calling session over input files (gsnp_job_*): parse and calibrate
    once, then repeat calling passes on device-resident inputs."""

    def __init__(self, fasta=None, soap=None, dbsnp=None, matrix_in=None, synth=None, **kw):
        """From files, or with synth=(config, seed, scale) from a seeded data
        set generated in memory (gsnp_job_synth)."""
        L = lib()
        check(L.gsnp_device_init(0))
        self.h = None
        self._p = params(**kw)
        h = ctypes.c_void_p()
        if synth is not None:
            rl = ctypes.c_int()
            cfg, seed, scale = synth
            check(L.gsnp_job_synth(int(cfg), int(seed), float(scale), ctypes.byref(self._p), _b(matrix_in),
                                   ctypes.byref(h), ctypes.byref(rl)))
            self.read_len = rl.value
        else:
            check(L.gsnp_job_open(_b(fasta), _b(soap), _b(dbsnp), _b(matrix_in), ctypes.byref(self._p),
                                  ctypes.byref(h)))
            self.read_len = self._p.read_length
        self.h = h

    def __len__(self) -> int:
        return lib().gsnp_job_count(self.h)

    def info(self, i: int) -> tuple[str, int, int]:
        name = ctypes.create_string_buffer(1024)
        n = ctypes.c_uint32()
        r = ctypes.c_uint64()
        check(lib().gsnp_job_info(self.h, i, name, 1024, ctypes.byref(n), ctypes.byref(r)))
        return name.value.decode(), n.value, r.value

    def select(self, i: int) -> None:
        check(lib().gsnp_job_select(self.h, i))
        self._length = self.info(i)[1]

    def call(self, batched: int = 0, emit_text: int = 1) -> int:
        """A calling pass over the selected chromosome; emit_text 0 no rows,
        1 rows in device memory, 2 rows copied to host memory.  Returns the row bytes."""
        nb = ctypes.c_uint64()
        check(lib().gsnp_job_call(self.h, int(batched), int(emit_text), ctypes.byref(nb)))
        return nb.value

    def set_regions(self, path) -> None:
        """-T target regions for later calls and writes (gsnp_job_set_regions); None clears them."""
        check(lib().gsnp_job_set_regions(self.h, _b(path)))

    def write(self, out, compress: bool = False, batched: int = 0) -> None:
        """Every chromosome called and written as .cns text or -C files (gsnp_job_write)."""
        check(lib().gsnp_job_write(self.h, _b(out), int(bool(compress)), int(batched)))

    def run_pass(self, batched: int = -1, emit_text: int = 2) -> tuple[int, int]:
        """A calling pass over every chromosome (gsnp_job_pass); batched < 0:
        one span per chromosome, 0: the per-window path.  Returns (row bytes, sites)."""
        nb, ns = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().gsnp_job_pass(self.h, int(batched), int(emit_text), ctypes.byref(nb), ctypes.byref(ns)))
        return nb.value, ns.value

    def columns(self, batched: int = 0) -> int:
        """A calling pass whose -C chunks are encoded on the device and copied to
        host memory without writing files (gsnp_job_columns); returns their bytes."""
        nb = ctypes.c_uint64()
        check(lib().gsnp_job_columns(self.h, int(batched), ctypes.byref(nb)))
        return nb.value

    def column_chunks(self, batched: int = 0):
        """A calling pass over the selected chromosome handing back every
        window's -C chunks (gsnp_job_columns_to): a list of
        (chromosome, 1-based start, sites, [five chunk byte strings])."""
        out = []

        def take(user, chr_, start, n, chunk, nbytes):
            try:
                out.append((chr_.decode(), start, n, [ctypes.string_at(chunk[f], nbytes[f]) for f in range(5)]))
                return 0
            except Exception:  # an exception must not cross the C frame
                return 1

        cb = CHUNK_FN(take)
        nb = ctypes.c_uint64()
        check(lib().gsnp_job_columns_to(self.h, int(batched), cb, None, ctypes.byref(nb)))
        return out

    @property
    def ctx(self):
        return lib().gsnp_job_context(self.h)

    def profile(self, enable: bool) -> dict:
        ms = (ctypes.c_double * len(STAGES))()
        cnt = (ctypes.c_uint64 * len(STAGES))()
        check(lib().gsnp_profile(self.ctx, int(enable), ms, cnt, len(STAGES)))
        return {s: (ms[i], cnt[i]) for i, s in enumerate(STAGES)}

    def autotune(self, start: int = 0, n_sites: int | None = None, reps: int = 2) -> dict:
        """Tune the fused launch of the batched form on the selected chromosome
        (gsnp_autotune); returns the chosen shape."""
        if n_sites is None:
            n_sites = self._length - start
        cfg = (ctypes.c_uint32 * 13)()
        check(lib().gsnp_autotune(self.ctx, start, n_sites, reps, cfg))
        return {"smallest_network": 8 << cfg[0], "waves_per_block": list(cfg[1:7]), "grid_cap": list(cfg[7:13])}

    def counters(self, reset: bool = True) -> dict:
        """Records called, sites finished on the host and rank-sum normal
        approximations finished on the device since the last reset (gsnp_counters)."""
        v = (ctypes.c_uint64 * 4)()
        check(lib().gsnp_counters(self.ctx, v, 4, int(reset)))
        return {"records": v[0], "host_fixups": v[1], "host_fixups_normal": v[2], "device_normal": v[3]}

    def records(self, reset: bool = True) -> int:
        v = ctypes.c_uint64()
        check(lib().gsnp_stats(self.ctx, ctypes.byref(v), int(reset)))
        return v.value

    def close(self) -> None:
        if self.h:
            lib().gsnp_job_close(self.h)
            self.h = None

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.gsnp_job_close(self.h)
            self.h = None

