"""Adjoint of the calling likelihood (gsnp_adjoint_*) on a seeded slice of the
miscalibrated 2 Mb workload: determinism, agreement with a torch-autograd
statement of the same objective, central finite differences, agreement with
the calling path's own likelihoods, and a short refit."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def adj_setup(built, tmp_path_factory):
    import gsnp_amd
    from gsnp_amd.adjoint import Adjoint

    root = tmp_path_factory.mktemp("adjoint")
    _, rl = gsnp_amd.synth(4, 21, 0.02, root / "c4")  # 40 kb, 30x, 100 bp
    fa, so = root / "c4.fa", root / "c4.soap"
    p0 = gsnp_amd.matrix_gen(fa, so, read_length=rl)
    job = gsnp_amd.Job(fa, so, read_length=rl)
    adj = Adjoint(job)
    log_het = math.log(job._p.het_novel_r)
    yield dict(job=job, adj=adj, p0=p0, log_het=log_het, root=root, fa=fa, so=so, rl=rl)
    job.close()


def test_adjoint_is_deterministic(adj_setup):
    adj, p0, lh = adj_setup["adj"], adj_setup["p0"], adj_setup["log_het"]
    a = adj.eval(p0, lh)
    b = adj.eval(p0, lh)
    assert np.float64(a[0]).tobytes() == np.float64(b[0]).tobytes()
    assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
    assert np.float64(a[2]).tobytes() == np.float64(b[2]).tobytes()
    assert adj.n_records > 1000 and np.count_nonzero(a[1]) > 100


def test_adjoint_matches_torch_autograd(adj_setup):
    import torch

    from gsnp_amd.adjoint import reference_loglik_torch

    adj, p0, lh = adj_setup["adj"], adj_setup["p0"], adj_setup["log_het"]
    ll, grad, dh = adj.eval(p0, lh)
    keys, off, prow = adj.export()
    lp, mask = adj.priors(lh)
    log_p = torch.tensor(np.log(p0.reshape(256, 256, 4, 4)[:64]), dtype=torch.float64, requires_grad=True)
    log_het = torch.tensor(lh, dtype=torch.float64, requires_grad=True)
    ref = reference_loglik_torch(log_p, log_het, keys, off, prow, lp, mask, lh)
    ref.backward()
    assert abs(float(ref) - ll) <= 1e-11 * abs(ll)
    g_ref = log_p.grad.numpy()
    scale = np.abs(g_ref).max()
    assert np.allclose(grad, g_ref, rtol=1e-9, atol=1e-12 * scale), float(np.abs(grad - g_ref).max())
    assert abs(float(log_het.grad) - dh) <= 1e-9 * abs(dh)


def test_adjoint_central_differences(adj_setup):
    from gsnp_amd.adjoint import central_differences

    adj, p0, lh = adj_setup["adj"], adj_setup["p0"], adj_setup["log_het"]
    _, grad, dh = adj.eval(p0, lh)
    flat = grad.reshape(-1)
    nz = np.flatnonzero(np.abs(flat) > 1e-3 * np.abs(flat).max())
    rng = np.random.default_rng(7)
    pick = rng.choice(nz, size=min(24, nz.size), replace=False)
    # flat gradient index (q<64 rows) and flat p_matrix index coincide: q<<12 | c<<4 | a<<2 | o
    fd, fd_het = central_differences(adj, p0, lh, pick, h=1e-6)
    rel = np.abs(fd - flat[pick]) / np.abs(flat[pick])
    assert rel.max() <= 1e-6, (rel.max(), pick[np.argmax(rel)])
    assert abs(fd_het - dh) <= 1e-6 * abs(dh)
    # an entry no record uses has a zero gradient and a zero difference
    zero = np.flatnonzero(flat == 0)[:2]
    fz, _ = central_differences(adj, p0, lh, zero, h=1e-6)
    assert np.all(fz == 0.0)


def test_adjoint_objective_uses_the_calling_likelihoods(adj_setup):
    """With the calibration table, the adjoint's objective equals the one
    computed from the per-window calling path's own log-likelihoods."""
    import gsnp_amd

    job, adj, p0, lh = adj_setup["job"], adj_setup["adj"], adj_setup["p0"], adj_setup["log_het"]
    site_terms = adj.eval(p0, lh, site_terms=True)[3]
    _, length, _ = job.info(0)
    assert length < gsnp_amd.WIN  # one window
    job.call(batched=0, emit_text=False)
    lk = np.zeros((length, 10))
    out = gsnp_amd.WindowOut()
    out.lk10 = lk.ctypes.data
    gsnp_amd.check(gsnp_amd.lib().gsnp_window_results(job.ctx, length, ctypes.byref(out)))
    _, _, prow = adj.export()
    lp, _ = adj.priors(lh)
    called = prow != 0xFFFFFFFF
    v = (lp[np.where(called, prow, 0)] + lk) * math.log(10.0)
    m = v.max(axis=1, keepdims=True)
    ll_sites = (m[:, 0] + np.log(np.exp(v - m).sum(axis=1)))
    assert np.allclose(ll_sites[called], site_terms[called], rtol=1e-12, atol=1e-9)
    assert np.all(site_terms[~called] == 0.0)


def test_refit_raises_the_objective(adj_setup):
    from gsnp_amd.adjoint import refit

    adj, p0, lh = adj_setup["adj"], adj_setup["p0"], adj_setup["log_het"]
    p1, trace = refit(adj, p0, lh, steps=5, lr=0.02)
    assert trace[-1] > trace[0]
    rows = p1.reshape(256, 256, 4, 4)[:64, :100]
    assert np.allclose(rows.sum(axis=-1), 1.0, atol=1e-12)


def test_refit_moves_toward_the_generator_truth(built, tmp_path):
    """On the miscalibrated workload (qualities 3 above the truth over q30 and
    a cycle offset up to 2, errors drawn independently per base), a refit
    under the calling model that matches the generator (-p 1 -g 1: no
    dependency adjustment) moves the table toward the generator's error
    rates 10^(-(q - delta)/10): its RMS log10 distance over q 10-40 falls
    clearly below the counting-trained table's, whose miscalibration it
    inherits from the stated qualities (0.278 -> 0.185 on this 200 kb
    slice; 0.269 -> 0.085 on the 2 Mb workload, where the rows hold ten
    times the bases)."""
    import gsnp_amd
    from gsnp_amd.adjoint import Adjoint, refit, truth_distance

    _, rl = gsnp_amd.synth(4, 21, 0.1, tmp_path / "c4")  # 200 kb, 30x, 100 bp
    fa, so = tmp_path / "c4.fa", tmp_path / "c4.soap"
    p0 = gsnp_amd.matrix_gen(fa, so, read_length=rl)
    job = gsnp_amd.Job(fa, so, read_length=rl, pcr_dependency=1.0, global_dependency=1.0)
    try:
        adj = Adjoint(job)
        p1, trace = refit(adj, p0, math.log(job._p.het_novel_r), steps=50)
    finally:
        job.close()
    d0, d1 = truth_distance(p0, rl), truth_distance(p1, rl)
    print(f"[adjoint] truth distance: counting {d0:.4f}, refit {d1:.4f}", flush=True)
    assert trace[-1] > trace[0]
    assert d1 < 0.75 * d0, (d0, d1)
