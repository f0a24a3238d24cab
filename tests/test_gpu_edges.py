"""Data sets with only one kind of datum (MI355X only): V2 only (nbis = 0, no
triangle legs in the adjoint coefficients or the line search) and T3 only (npow = 0,
every uv point reached through the triangles), sliced from the 64^2 workload.  The
objective and gradient in FP32 and BF16x3 against the FP64 oracle with the bounds of
test_gpu_parity.py; the accelerated loop against orc_pcg (FP32, the same rules:
equal DFT and batch counts); the grouped batch over a mix of them bit-identical to
the single runs; a data set with neither, and an unsupported width, are refused."""
import dataclasses

import numpy as np
import pytest

from gpair_amd import synth
from gpair_amd.api import GpairContext, GpairError
from oracle import Oracle

pytestmark = pytest.mark.gpu

ALPHA = 100.0
# max relative errors (visibilities, chi2, data gradient) per precision
BOUNDS = {0: (2e-6, 1e-8, 2e-6), 1: (8e-6, 2e-7, 4e-5)}


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def v2_only(p):
    return dataclasses.replace(p, nbis=0, data=p.data[:p.npow].copy(), err=p.err[:p.npow].copy(),
                               phasor=p.phasor[:0].copy(), uvpnt=p.uvpnt[:0].copy(), sign=p.sign[:0].copy())


def t3_only(p):
    return dataclasses.replace(p, npow=0, data=p.data[p.npow:].copy(), err=p.err[p.npow:].copy())


KINDS = {"v2": v2_only, "t3": t3_only}


@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("kind", list(KINDS))
def test_objective_and_gradient_one_kind(kind, prec, orc):
    prob = KINDS[kind](synth.make_problem("cfg1", seed=1))
    print(f"case {kind} prec {prec}: nuv={prob.nuv} npow={prob.npow} nbis={prob.nbis}", flush=True)
    P = orc.unpack(prob)
    rng = np.random.default_rng(3)
    img = (prob.start.astype(np.float64) * (1.0 + rng.random(prob.W * prob.W))).astype(np.float32)
    with GpairContext.from_problem(prob, image=img) as ctx:
        ctx.set_precision(prec)
        chi2 = ctx.gpu_get_chi2_curr()
        vis = ctx.vis().astype(np.complex128)
        ctx.gpu_compute_data_gradient_curr()
        gd = ctx.get("dg").astype(np.float64)
    J, chi2_o, S_o, vis_o, mock_o, F = orc.objective(P, img.astype(np.float64), ALPHA)
    g_o, gd_o, ge_o = orc.gradient(P, img.astype(np.float64), ALPHA, vis_o, mock_o)
    ev = np.abs(vis - vis_o).max() / np.abs(vis_o).max()
    ec = abs(chi2 - chi2_o) / chi2_o
    eg = np.abs(gd - gd_o).max() / np.abs(gd_o).max()
    print(f"  vis {ev:.2e} chi2 {ec:.2e} grad {eg:.2e}", flush=True)
    bv, bc, bg = BOUNDS[prec]
    assert ev <= bv and ec <= bc and eg <= bg


@pytest.mark.parametrize("kind", list(KINDS))
def test_pcg_loop_one_kind(kind, orc):
    prob = KINDS[kind](synth.make_problem("cfg1", seed=1))
    print(f"case {kind}: npow={prob.npow} nbis={prob.nbis}", flush=True)
    P = orc.unpack(prob)
    alpha = prob.meta["alpha"]
    nit = 10
    with GpairContext.from_problem(prob) as ctx:
        res, tr = ctx.pcg_run(alpha, precond=1, max_iter=nit, chi2r_target=0.0)
    img_o, tr_o = orc.pcg(P, prob.start.astype(np.float64), alpha, precond=1, max_iter=nit, chi2r_target=0.0)
    rel = np.abs(tr[:, 0] - tr_o[:, 0]) / np.abs(tr_o[:, 0])
    print("  trajectory rel diff", rel, flush=True)
    assert res.iterations == nit and len(tr_o) == nit
    assert rel.max() <= 2e-4   # the FP32 bound of the full data set's loop test
    assert np.array_equal(tr[:, 5:8], tr_o[:, 5:8])


def test_batch_with_one_kind_problems():
    from gpair_amd.api import batch_pcg_run
    base = synth.make_problem("cfg1", seed=2)
    probs = [v2_only(base), t3_only(base), base]
    fields = ("iterations", "converged", "forward_dfts", "adjoint_dfts", "ls_batches", "J")
    single = []
    for p in probs:
        with GpairContext.from_problem(p) as ctx:
            ctx.set_precision(1)
            r, _ = ctx.pcg_run(p.meta["alpha"], precond=1, max_iter=60, chi2r_target=1.0)
            single.append((tuple(getattr(r, f) for f in fields), ctx.get("ci")))
    ctxs = [GpairContext.from_problem(p) for p in probs]
    try:
        for c in ctxs:
            c.set_precision(1)
        res = batch_pcg_run(ctxs, [p.meta["alpha"] for p in probs], max_iter=60, chi2r_target=1.0)
        for (rs, img), r, c in zip(single, res, ctxs):
            assert tuple(getattr(r, f) for f in fields) == rs
            assert np.array_equal(c.get("ci"), img)
    finally:
        for c in ctxs:
            c.close()


def test_no_data_is_refused():
    p = synth.make_problem("cfg1", seed=1)
    ctx = GpairContext(0)
    try:
        with pytest.raises(GpairError, match="no V2 and no T3"):
            ctx.gpu_copy_data(p.data[:0], p.err[:0], p.nuv, p.phasor[:0], 0, p.uvpnt4()[:0], p.sign4()[:0],
                              p.model, p.W)
    finally:
        ctx.close()


@pytest.mark.parametrize("W", [32, 100, 320, 1024])
def test_unsupported_width_is_refused(W):
    """widths outside 64, 128, 192, 256, 512 fail with an error, before any launch."""
    p = synth.make_problem("cfg1", seed=1)
    ctx = GpairContext(0)
    try:
        with pytest.raises(GpairError, match="width"):
            ctx.gpu_copy_data(p.data, p.err, p.nuv, p.phasor, p.npow, p.uvpnt4(), p.sign4(),
                              np.ones(W * W, np.float32), W)
    finally:
        ctx.close()
