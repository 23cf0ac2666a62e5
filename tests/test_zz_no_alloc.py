"""gsnp_sort_likelihood_noalloc (caller-owned device memory, no allocation)
against gsnp_sort_likelihood on the same records: sorted rows and the ten
log-likelihoods of every site equal bit for bit, on two calls in a row (the
second smaller, so stale list entries of the first would show), plus its
argument checks."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_MEMCPY_H2D, HIP_MEMCPY_D2H = 1, 2


def _hip():
    """The HIP runtime that libgsnp.so itself loaded (same soname), for the
    caller-side buffers of the test."""
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevArray:
    def __init__(self, hip, nbytes: int):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(1, nbytes)) == 0
        self.nbytes = nbytes

    def put(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, HIP_MEMCPY_H2D) == 0

    def get(self, a: np.ndarray) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, HIP_MEMCPY_D2H) == 0
        return a

    def free(self) -> None:
        if self.p:
            self.hip.hipFree(self.p)
            self.p = None


def _context(gsnp_amd):
    L = gsnp_amd.lib()
    p = gsnp_amd.params(read_length=100)
    pm = np.full(256 * 256 * 16, 1.0)
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
    return ctx


def _records(rng, n_sites: int):
    counts = rng.integers(0, 257, n_sites).astype(np.int32)
    edge = [0, 1, 2, 8, 9, 16, 17, 32, 33, 64, 65, 128, 129, 255, 256, 7, 15, 31, 63, 127]
    counts[: len(edge)] = edge[:n_sites]
    recs = np.full((n_sites, 256), 0xFFFFFFFF, dtype=np.uint32)
    for i in range(n_sites):
        n = int(counts[i])
        base = rng.integers(0, 4, n)
        qs = rng.integers(1, 41, n)
        cyc = rng.integers(0, 100, n) if i % 3 else rng.integers(0, 4, n)  # duplicate keys too
        st = rng.integers(0, 2, n)
        recs[i, :n] = (base << 23) | ((64 - qs) << 17) | (cyc << 9) | (st << 8) | 1
    return recs, counts


def test_noalloc_matches_sort_likelihood(built):
    import gsnp_amd

    L = gsnp_amd.lib()
    hip = _hip()
    ctx = _context(gsnp_amd)
    rng = np.random.default_rng(77)
    cap = 3000
    d_rec = DevArray(hip, cap * 256 * 4)
    d_cnt = DevArray(hip, cap * 4)
    d_lk = DevArray(hip, cap * 80)
    try:
        for n_sites in (3000, 1100):  # second call smaller than the first
            print(f"[no-alloc] {n_sites} sites", flush=True)
            recs, counts = _records(rng, n_sites)
            want_rec = recs.copy()
            want_lk = np.zeros((n_sites, 10))
            gsnp_amd.check(L.gsnp_sort_likelihood(ctx, want_rec.ctypes.data, 256, counts.ctypes.data, n_sites,
                                                  want_lk.ctypes.data))
            d_rec.put(recs)
            d_cnt.put(counts)
            d_lk.put(np.full(cap * 10, np.nan))
            gsnp_amd.check(L.gsnp_sort_likelihood_noalloc(ctx, d_rec.p, d_cnt.p, n_sites, d_lk.p))
            got_rec = d_rec.get(np.zeros((n_sites, 256), dtype=np.uint32))
            got_lk = d_lk.get(np.zeros((cap, 10)))
            for i in range(n_sites):
                n = int(counts[i])
                assert np.array_equal(got_rec[i, :n], want_rec[i, :n]), f"site {i} (n={n}) records differ"
                assert np.array_equal(got_rec[i, n:], recs[i, n:]), f"site {i}: slots past the count were written"
            assert np.array_equal(got_lk[:n_sites].view(np.uint64), want_lk.view(np.uint64)), "likelihoods differ"
            assert np.isnan(got_lk[n_sites:]).all(), "likelihoods written past n_sites"
    finally:
        d_rec.free()
        d_cnt.free()
        d_lk.free()
        L.gsnp_destroy(ctx)


def test_noalloc_rejects_bad_arguments(built):
    import gsnp_amd

    L = gsnp_amd.lib()
    hip = _hip()
    ctx = _context(gsnp_amd)
    d_rec = DevArray(hip, 4 * 256 * 4)
    d_cnt = DevArray(hip, 4 * 4)
    d_lk = DevArray(hip, 4 * 80)
    try:
        # more sites than the fixed working array holds: refused before any device work
        rc = L.gsnp_sort_likelihood_noalloc(ctx, d_rec.p, d_cnt.p, 262145, d_lk.p)
        assert rc == 8, rc
        assert b"262144" in L.gsnp_last_error()
        # a count above 256 slots: refused by the device check before the sort
        rec = np.full((4, 256), (1 << 23) | ((64 - 30) << 17) | (5 << 9) | 1, dtype=np.uint32)
        d_rec.put(rec)
        d_cnt.put(np.array([3, 257, 0, 1], dtype=np.int32))
        rc = L.gsnp_sort_likelihood_noalloc(ctx, d_rec.p, d_cnt.p, 4, d_lk.p)
        assert rc == 8, rc
        # and a valid call afterwards still works (the error word is cleared)
        d_cnt.put(np.array([3, 2, 0, 1], dtype=np.int32))
        gsnp_amd.check(L.gsnp_sort_likelihood_noalloc(ctx, d_rec.p, d_cnt.p, 4, d_lk.p))
        assert L.gsnp_sort_likelihood_noalloc(None, d_rec.p, d_cnt.p, 4, d_lk.p) == 8
    finally:
        d_rec.free()
        d_cnt.free()
        d_lk.free()
        L.gsnp_destroy(ctx)
