"""Data I/O around the HIP path (MI355X only): a reconstruction set up from the
reference's own OIFITS file, checkpoints of the accelerated loop (a resumed run
equals the uninterrupted one bit for bit) and the loops' periodic snapshots."""
import os

import numpy as np
import pytest

from gpair_amd import oifits, synth
from gpair_amd.api import GpairContext, GpairError
from oracle import Oracle

pytestmark = pytest.mark.gpu

# the reference's own data file (gpair1/trunk/2004contest1.oifits), kept as a fixture
REF_FILE = os.path.join(os.path.dirname(__file__), "golden", "2004contest1.oifits")


def test_reconstruction_from_the_reference_oifits_file():
    """the reference's set-up (gpair1/trunk/gpair.c:78-229 defaults: 128 px of
    0.15 mas, Gaussian prior of 5 mas and flux 10, alpha 1000) from its data file:
    the objective of the start image against the FP64 oracle, then the loop lowers
    it"""
    prob = oifits.problem(REF_FILE, 128, 0.15, model_type=3, model_width=5.0, model_flux=10.0)
    print(f"contest file: {prob.npow} V2, {prob.nbis} T3, nuv {prob.nuv}", flush=True)
    orc = Oracle()
    P = orc.unpack(prob)
    J_o, chi2_o, S_o, *_ = orc.objective(P, prob.start.astype(np.float64), 1000.0)
    with GpairContext.from_problem(prob) as ctx:
        chi2 = ctx.gpu_get_chi2_curr()
        S = ctx.gpu_get_entropy_curr()
        r, tr = ctx.pcg_run(1000.0, max_iter=200, chi2r_target=0.0)
    print(f"  chi2 {chi2:.6e} (oracle {chi2_o:.6e}), J {tr[0, 0]:.6e} -> {r.J:.6e} in {r.iterations} its, "
          f"chi2r {r.chi2r:.3f}", flush=True)
    assert abs(chi2 - chi2_o) <= 2e-5 * chi2_o and abs(S - S_o) <= 2e-5 * abs(S_o)
    assert np.all(np.diff(tr[:, 0]) <= 1e-9 * np.abs(tr[:-1, 0])) and r.J < 0.5 * tr[0, 0]


@pytest.mark.parametrize("prec,graphs", [(0, False), (1, True)], ids=["fp32", "bf16x3-graphs"])
def test_checkpoint_resume_is_bit_identical(tmp_path, prec, graphs):
    prob = synth.make_problem("cfg1", seed=2, W=128, ntel=7, nch=6, nsnap=5)
    alpha = prob.meta["alpha"]
    N, M = 25, 30
    ck = str(tmp_path / "ck.fits")
    with GpairContext.from_problem(prob) as a:
        a.set_precision(prec)
        a.set_graphs(graphs)
        ra, tra = a.pcg_run(alpha, max_iter=N + M, chi2r_target=0.0)
        img_a = a.get("ci")
    with GpairContext.from_problem(prob) as b:
        b.set_precision(prec)
        b.set_graphs(graphs)
        rb, _ = b.pcg_run_checkpointed(alpha, max_iter=N, chi2r_target=0.0, save_to=ck)
        img_b = b.get("ci")
    assert rb.iterations == N and np.array_equal(oifits.read_image(ck).ravel(), img_b)
    with GpairContext.from_problem(prob) as c:   # a fresh context, start image untouched
        c.set_precision(prec)
        c.set_graphs(not graphs)                 # graphs or launches: the same bits
        rc, trc = c.pcg_run_checkpointed(alpha, max_iter=N + M, chi2r_target=0.0, resume_from=ck)
        img_c = c.get("ci")
        with pytest.raises(GpairError, match="alpha"):
            c.pcg_run_checkpointed(2 * alpha, max_iter=N + M, resume_from=ck)
    print(f"  J {ra.J!r} resumed {rc.J!r}; dfts {ra.forward_dfts}/{rc.forward_dfts}", flush=True)
    assert rc.J == ra.J and rc.chi2 == ra.chi2 and rc.entropy == ra.entropy
    assert np.array_equal(img_c, img_a)
    assert (rc.iterations, rc.forward_dfts, rc.adjoint_dfts, rc.ls_batches, rc.resyncs) == \
        (ra.iterations, ra.forward_dfts, ra.adjoint_dfts, ra.ls_batches, ra.resyncs)
    assert np.array_equal(trc[N:], tra[N:])


def test_snapshots_of_the_loops(tmp_path):
    prob = synth.make_problem("cfg1", seed=3)
    alpha = prob.meta["alpha"]
    snap = str(tmp_path / "reconst.fits")
    with GpairContext.from_problem(prob) as ctx:
        ctx.set_snapshot(snap, every=1, pixsize=prob.pixellation)
        ctx.pcg_run(alpha, max_iter=3, chi2r_target=0.0)
    with GpairContext.from_problem(prob) as ref:
        ref.pcg_run(alpha, max_iter=2, chi2r_target=0.0)
        after2 = ref.get("ci")
    # the last snapshot of a 3-iteration run is the image after two commits
    assert np.array_equal(oifits.read_image(snap).ravel(), after2)
    snap2 = str(tmp_path / "cg.fits")
    with GpairContext.from_problem(prob) as ctx:
        ctx.set_snapshot(snap2, every=5)
        ctx.cg_run(alpha, method=2, max_iter=7, chi2r_target=0.0)
        ctx.set_snapshot(None)
    img = oifits.read_image(snap2)
    assert img.shape == (prob.W, prob.W) and np.all(np.isfinite(img)) and img.min() > 0


@pytest.mark.parametrize("saved_factor,context_factor", [(True, False), (False, True)],
                         ids=["ck-factor", "context-factor"])
def test_checkpoint_with_factor_and_graphs(tmp_path, saved_factor, context_factor):
    """the gradient-phase graphs hold the preconditioner factor's pointer: a resume
    that restores a factor (or drops the context's) captures them after the
    checkpoint is loaded, so the resumed graph run equals the uninterrupted run
    without graphs bit for bit"""
    prob = synth.make_problem("cfg1", seed=4, W=128, ntel=7, nch=6, nsnap=5)
    alpha = prob.meta["alpha"]
    rng = np.random.default_rng(3)
    pf = (0.5 + rng.random(prob.W ** 2)).astype(np.float32)
    other = (0.5 + rng.random(prob.W ** 2)).astype(np.float32)
    N, M = 15, 20
    ck = str(tmp_path / "ck.fits")
    with GpairContext.from_problem(prob) as a:          # uninterrupted, launched
        a.set_precision(1)
        if saved_factor:
            a.set_pfactor(pf)
        ra, tra = a.pcg_run(alpha, max_iter=N + M, chi2r_target=0.0)
        img_a = a.get("ci")
    with GpairContext.from_problem(prob) as b:          # first part, graphs
        b.set_precision(1)
        b.set_graphs(True)
        if saved_factor:
            b.set_pfactor(pf)
        b.pcg_run_checkpointed(alpha, max_iter=N, chi2r_target=0.0, save_to=ck)
    with GpairContext.from_problem(prob) as c:          # resumed, graphs, another factor state
        c.set_precision(1)
        c.set_graphs(True)
        if context_factor:
            c.set_pfactor(other)
        rc, trc = c.pcg_run_checkpointed(alpha, max_iter=N + M, chi2r_target=0.0, resume_from=ck)
        img_c = c.get("ci")
    print(f"  J {ra.J!r} resumed {rc.J!r}", flush=True)
    assert rc.J == ra.J and np.array_equal(img_c, img_a) and np.array_equal(trc[N:], tra[N:])
