"""Adjoint of the calling likelihood, its autograd wrapper and the table refit.

For the sites of a span of the selected chromosome, the objective is

    l(ln p, ln het) = sum over called sites s of ln sum_g prior_s(g) 10^{lk_s(g)}

with lk_s the genotype log10-likelihoods of the calling path (records,
dependency-adjusted qualities and genotype order as in
kernel/likelihood_kernel.cu:137-256) and prior_s the calling priors
(cpp/prior.cc).  The value and its gradient with respect to ln p_matrix
(q < 64) and ln het are computed by the library's HIP kernels
(gsnp_adjoint_prepare / gsnp_adjoint_eval), with a deterministic
sort-then-segmented-reduce accumulation.  This module is the Python side:

* ``Adjoint``       -- a prepared span; ``eval`` returns (l, dl/dln p, dl/dln het).
* ``LogMarginal``   -- torch.autograd.Function over (ln p, ln het).
* ``central_differences`` -- the finite-difference check of the gradient.
* ``refit``         -- maximum-marginal-likelihood refit of the table,
  parameterised by a softmax over the observed base so every (q, cycle, ref)
  row stays a distribution; the plain calling path then judges which calls
  change (``calls_changed``).

The reference has no gradient; nothing here has a reference counterpart.
"""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

import numpy as np

from . import _b, check, lib

PM_SHAPE = (256, 256, 4, 4)  # p_matrix[q][cycle][ref][obs], index q<<12 | cycle<<4 | ref<<2 | obs
GRAD_SHAPE = (64, 256, 4, 4)
GENOTYPES = [(a1, a2) for a1 in range(4) for a2 in range(a1, 4)]  # AA AC AT AG CC CT CG TT TG GG
NO_ROW = 0xFFFFFFFF


class Adjoint:
    """gsnp_adjoint_* over sites [start, start + n_sites) of chromosome `index` of a Job."""

    def __init__(self, job, index: int = 0, start: int = 0, n_sites: int | None = None):
        _, length, _ = job.info(index)
        job.select(index)
        if n_sites is None:
            n_sites = length - start
        self.job = job
        self.ctx = job.ctx
        check(lib().gsnp_adjoint_prepare(self.ctx, start, n_sites))
        ns, nr, rows = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32()
        check(lib().gsnp_adjoint_info(self.ctx, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(rows)))
        self.n_sites, self.n_records, self.n_prior_rows = ns.value, nr.value, rows.value

    def eval(self, p_matrix: np.ndarray, log_het: float, site_terms: bool = False):
        """(l, dl/dln p [64,256,4,4], dl/dln het[, per-site l terms])."""
        pm = np.ascontiguousarray(p_matrix, dtype=np.float64).reshape(-1)
        if pm.size != 256 * 256 * 16:
            raise ValueError("p_matrix must hold 256*256*16 doubles")
        grad = np.empty(64 * 256 * 16, dtype=np.float64)
        ll, dh = ctypes.c_double(), ctypes.c_double()
        sites = np.empty(max(1, self.n_sites), dtype=np.float64) if site_terms else None
        check(lib().gsnp_adjoint_eval(self.ctx, pm.ctypes.data, float(log_het), ctypes.byref(ll), grad.ctypes.data,
                                      ctypes.byref(dh), sites.ctypes.data if sites is not None else None))
        out = (ll.value, grad.reshape(GRAD_SHAPE), dh.value)
        return out + (sites[: self.n_sites],) if site_terms else out

    def export(self):
        """(keys per record, offsets per site, prior row per site) of the prepared span."""
        keys = np.empty(max(1, self.n_records), dtype=np.uint32)
        off = np.empty(self.n_sites + 1, dtype=np.uint64)
        prow = np.empty(max(1, self.n_sites), dtype=np.uint32)
        check(lib().gsnp_adjoint_export(self.ctx, keys.ctypes.data, off.ctypes.data, prow.ctypes.data))
        return keys[: self.n_records], off, prow[: self.n_sites]

    def priors(self, log_het: float):
        """(log10 prior rows [rows, 10], het masks [rows]) at ln het = log_het."""
        lp = np.empty((self.n_prior_rows, 10), dtype=np.float64)
        mask = np.empty(self.n_prior_rows, dtype=np.uint32)
        check(lib().gsnp_adjoint_priors(self.ctx, float(log_het), lp.ctypes.data, mask.ctypes.data,
                                        self.n_prior_rows))
        return lp, mask


def _torch_function():
    import torch

    class LogMarginal(torch.autograd.Function):
        """l(ln p, ln het) of an Adjoint span.  ln p is [64, 256, 4, 4] float64 (the
        q < 64 rows of the table); rows q >= 64 come from `base` and are unused."""

        @staticmethod
        def forward(ctx, log_p, log_het, adj: Adjoint, base: np.ndarray):
            pm = np.array(base, dtype=np.float64).reshape(PM_SHAPE)
            pm[:64] = np.exp(log_p.detach().cpu().numpy())
            ll, g, dh = adj.eval(pm, float(log_het))
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(g)).to(log_p.device),
                                  torch.tensor(dh, dtype=torch.float64, device=log_p.device))
            return torch.tensor(ll, dtype=torch.float64, device=log_p.device)

        @staticmethod
        def backward(ctx, grad_out):
            g, dh = ctx.saved_tensors
            return grad_out * g, grad_out * dh, None, None

    return LogMarginal


_LOGMARGINAL = None


def log_marginal(log_p, log_het, adj: Adjoint, base: np.ndarray):
    """torch scalar l; differentiable in log_p ([64,256,4,4]) and log_het."""
    global _LOGMARGINAL
    if _LOGMARGINAL is None:
        _LOGMARGINAL = _torch_function()
    return _LOGMARGINAL.apply(log_p, log_het, adj, base)


def reference_loglik_torch(log_p, log_het, keys, offsets, prow, lp_rows, het_mask, log_het0):
    """The same objective written with torch ops (float64), for autograd checks.

    lp_rows / het_mask are the prior rows at ln het = log_het0; a masked
    genotype's prior is proportional to het, so its row moves by
    (log_het - log_het0) / ln 10."""
    import torch

    keys = torch.as_tensor(keys.astype(np.int64))
    q, cyc, o = keys >> 10, (keys >> 2) & 255, keys & 3
    p = torch.exp(log_p)
    pq = p[q, cyc, :, o]  # [R, 4]: p[q, cycle, ref=a, obs=o]
    a1 = torch.tensor([g[0] for g in GENOTYPES])
    a2 = torch.tensor([g[1] for g in GENOTYPES])
    score = torch.log10(0.5 * pq[:, a1] + 0.5 * pq[:, a2])  # [R, 10]
    counts = np.diff(offsets.astype(np.int64))
    seg = torch.as_tensor(np.repeat(np.arange(len(counts)), counts))
    lk = torch.zeros(len(counts), 10, dtype=torch.float64).index_add_(0, seg, score)
    called = torch.as_tensor(prow != NO_ROW)
    rows = torch.as_tensor(np.where(prow != NO_ROW, prow, 0).astype(np.int64))
    masks = torch.as_tensor(((het_mask[:, None] >> np.arange(10)) & 1).astype(np.float64))
    lpr = torch.as_tensor(lp_rows) + masks * (log_het - log_het0) / math.log(10.0)
    v = (lpr[rows] + lk) * math.log(10.0)
    return torch.logsumexp(v, dim=1)[called].sum()


def central_differences(adj: Adjoint, p_matrix: np.ndarray, log_het: float, entries, h: float = 1e-6):
    """dl/dln p at the given flat p_matrix entries and dl/dln het by central
    differences, differencing per site so unaffected sites cancel exactly."""
    pm = np.ascontiguousarray(p_matrix, dtype=np.float64).reshape(-1)
    out = []
    for e in entries:
        plus, minus = pm.copy(), pm.copy()
        plus[e] *= math.exp(h)
        minus[e] *= math.exp(-h)
        lp = adj.eval(plus, log_het, site_terms=True)[3]
        lm = adj.eval(minus, log_het, site_terms=True)[3]
        out.append(float(np.sum(lp - lm)) / (2 * h))
    lp = adj.eval(pm, log_het + h, site_terms=True)[3]
    lm = adj.eval(pm, log_het - h, site_terms=True)[3]
    return np.array(out), float(np.sum(lp - lm)) / (2 * h)


def write_matrix(path, p_matrix: np.ndarray, read_length: int) -> None:
    """The -M / -I calibration file (cpp/matrix.cc:215-241 layout: q, cycle, 16 values)."""
    pm = np.asarray(p_matrix, dtype=np.float64).reshape(256, 256, 16)
    with open(path, "w") as f:
        for q in range(64):
            for c in range(read_length):
                f.write(f"{q}\t{c}" + "".join(f"\t{v:.16e}" for v in pm[q, c]) + "\n")


def refit(adj: Adjoint, p0: np.ndarray, log_het: float, steps: int = 50, lr: float = 0.02):
    """Refit the q < 64 table rows by maximum marginal likelihood (Adam on a
    softmax parameterisation over the observed base).  Returns the refit
    p_matrix and the objective after each step."""
    import torch

    base = np.array(p0, dtype=np.float64).reshape(PM_SHAPE)
    theta = torch.tensor(np.log(np.maximum(base[:64], 1e-300)), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([theta], lr=lr)
    lh = torch.tensor(float(log_het), dtype=torch.float64)
    trace = []
    for _ in range(steps):
        opt.zero_grad()
        log_p = torch.log_softmax(theta, dim=-1)
        loss = -log_marginal(log_p, lh, adj, base)
        loss.backward()
        opt.step()
        trace.append(-float(loss))
    out = base.copy()
    out[:64] = torch.softmax(theta.detach(), dim=-1).numpy()
    return out.reshape(-1), trace


def truth_distance(p_matrix: np.ndarray, read_length: int, q_lo: int = 10, q_hi: int = 40) -> float:
    """Distance of a table from the error rates the miscalibrated workload was
    generated with: the RMS over rows (q, cycle), q_lo <= q <= q_hi and cycle
    < read_length, of log10 of the row's error rate (1 - p[q][cycle][t][t],
    averaged over the four reference bases t) minus log10 of the true
    10^(-(q - delta)/10), delta = 3 [q > 30] + 2 cycle / 99."""
    pm = np.asarray(p_matrix, dtype=np.float64).reshape(PM_SHAPE)
    d = []
    for q in range(q_lo, q_hi + 1):
        for c in range(read_length):
            err = float(np.mean([1.0 - pm[q, c, t, t] for t in range(4)]))
            true = 10.0 ** (-(q - (3.0 * (q > 30) + 2.0 * c / 99.0)) / 10.0)
            d.append(math.log10(max(err, 1e-300)) - math.log10(true))
    return float(np.sqrt(np.mean(np.square(d))))


def calls_changed(cns_a, cns_b) -> tuple[int, int]:
    """(rows whose genotype call differs, rows whose consensus quality differs)
    between two .cns files of the same sites."""
    ng = nq = 0
    with open(cns_a) as fa, open(cns_b) as fb:
        for la, lb in zip(fa, fb):
            ta, tb = la.split("\t", 5), lb.split("\t", 5)
            ng += ta[3] != tb[3]
            nq += ta[4] != tb[4]
    return ng, nq


def refit_and_judge(fasta, soap, workdir, steps: int = 50, lr: float = 0.02, **kw):
    """Counting-trained table -> refit -> plain calling with both tables; returns
    ((genotype calls changed, consensus qualities changed), objective trace, n_sites)."""
    from . import Job, matrix_gen, run

    work = Path(workdir)
    work.mkdir(parents=True, exist_ok=True)
    p0 = matrix_gen(fasta, soap, **kw)
    job = Job(fasta, soap, **kw)
    try:
        adj = Adjoint(job)
        log_het = math.log(job._p.het_novel_r)
        p1, trace = refit(adj, p0, log_het, steps=steps, lr=lr)
        n_sites = adj.n_sites
    finally:
        job.close()
    rl = kw.get("read_length", 100)
    m0, m1 = work / "p0.matrix", work / "p1.matrix"
    write_matrix(m0, p0, rl)
    write_matrix(m1, p1, rl)
    c0, c1 = work / "plain0.cns", work / "plain1.cns"
    run(fasta, soap, c0, matrix_in=m0, **kw)
    run(fasta, soap, c1, matrix_in=m1, **kw)
    return calls_changed(c0, c1), trace, n_sites
