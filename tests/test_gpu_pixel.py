"""Single-pixel moves (gpu_update_vis_fluxchange, gpair_update_vis_positionchange,
gpu_new_chi2, gpair_pixel_propose / apply / state) against the FP64 oracle
through the C ABI (MI355X only).

The pixel state's visibilities come from one forward DFT (FP32, or BF16x3) of the
image; every update after that is FP64.  Tolerances:
  trial visibilities   max|dV| <= 2e-6 max|V| (FP32 state), 8e-6 (BF16x3 state)
  gpu_new_chi2         relative 2e-5 (chi2 of the float trial visibilities)
  move scores dchi2    |d| <= 1e-3 |dchi2| + 1e-7 chi2 (FP64 differences from the DFT state)
  move scores dS       relative 1e-12 (FP64 from the float pixel values)
  move / inverse       |dchi2 + dchi2_inv| <= 1e-6 |dchi2| (a scored change of U is
                       rounded to float, an applied one is not)
  state chi2           relative 2e-5, flux relative 1e-12
"""
import numpy as np
import pytest

from gpair_amd import synth
from gpair_amd.api import PIXEL_MOVE, GpairContext, GpairError
from oracle import Oracle

pytestmark = pytest.mark.gpu

CASES = [("cfg1", dict(seed=1)), ("cfg1_w128", dict(seed=2, W=128, ntel=7, nch=6, nsnap=5)),
         ("cfg2", dict(seed=3))]
VIS_TOL = {0: 2e-6, 1: 8e-6}


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def _problem(name, kw):
    return synth.make_problem("cfg2" if name.startswith("cfg2") else "cfg1", **kw)


def _image(prob, seed):
    rng = np.random.default_rng(seed)
    return (prob.start.astype(np.float64) * (1.0 + rng.random(prob.W * prob.W))).astype(np.float32)


def _moves(img, W, K, seed):
    """K moves: one-pixel flux changes (x0.05 .. x20 of the pixel, or a bright
    point), every third a two-pixel move, one null move (k = 1)"""
    rng = np.random.default_rng(seed)
    n = W * W
    m = np.zeros(K, PIXEL_MOVE)
    m["pix_a"] = rng.choice(n, K, replace=False)
    m["pix_b"] = -1
    m["flux_a"] = img[m["pix_a"]] * np.exp(rng.uniform(np.log(0.05), np.log(20.0), K))
    two = np.arange(K) % 3 == 0
    pb = (m["pix_a"] + 1 + rng.integers(0, n - 1, K)) % n
    m["pix_b"][two] = pb[two]
    m["flux_b"][two] = img[pb[two]] * 2.5
    m["flux_a"][2] = img.sum() * 0.01
    m["flux_a"][1] = img[m["pix_a"][1]]
    m["pix_b"][1] = -1
    return m


def _oracle_moves(orc, P, img64, vis, F, m):
    return orc.px_propose(P, img64, vis, F, m["pix_a"], m["pix_b"], m["flux_a"].astype(np.float64),
                          m["flux_b"].astype(np.float64))


@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("name,kw", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_trial_visibilities_and_new_chi2(name, kw, prec, orc):
    prob = _problem(name, kw)
    print(f"case {name} precision {prec}: W={prob.W} nuv={prob.nuv}", flush=True)
    P = orc.unpack(prob)
    W = prob.W
    img = _image(prob, 7)
    img64 = img.astype(np.float64)
    vis_o, F_o = orc.image2vis(W, P["au"], P["av"], img64)
    with GpairContext.from_problem(prob, image=img) as ctx:
        ctx.set_precision(prec)
        x, y, new = 11, W // 2 + 3, float(img.sum() * 0.002)
        ctx.gpu_update_vis_fluxchange(x, y, new)
        tv = ctx.trial_vis().astype(np.complex128)
        chi2 = ctx.gpu_new_chi2()
        ref, _ = orc.px_update(P, vis_o, F_o, x + W * y, new - img64[x + W * y])
        ev = np.abs(tv - ref).max() / np.abs(ref).max()
        mock = orc.vis2data(P["npow"], P["nbis"], ref, P["uvpnt"], P["sign"], P["phasor"])
        chi2_o = orc.chi2(P["data"], P["err"], mock)
        ec = abs(chi2 - chi2_o) / chi2_o
        # position change: a quarter of pixel (5, 9)'s flux to (40, 21)
        f = float(img[5 + W * 9]) * 0.25
        ctx.update_vis_positionchange(5, 9, 40, 21, f)
        tv2 = ctx.trial_vis().astype(np.complex128)
        ref2, _ = orc.px_update(P, vis_o, F_o, [5 + W * 9, 40 + W * 21], [-f, f])
        ev2 = np.abs(tv2 - ref2).max() / np.abs(ref2).max()
        # the current image and its own chi2 are untouched
        assert np.array_equal(ctx.get("ci"), img)
    print(f"  flux change: vis {ev:.2e} chi2 {ec:.2e}; position change: vis {ev2:.2e}", flush=True)
    assert ev <= VIS_TOL[prec] and ev2 <= VIS_TOL[prec]
    assert ec <= 2e-5


@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_propose_scores_match_the_oracle(name, kw, prec, orc):
    prob = _problem(name, kw)
    print(f"case {name} precision {prec}: W={prob.W} nuv={prob.nuv} nbis={prob.nbis}", flush=True)
    P = orc.unpack(prob)
    W = prob.W
    img = _image(prob, 8)
    img64 = img.astype(np.float64)
    vis_o, F_o = orc.image2vis(W, P["au"], P["av"], img64)
    chi2_o = orc.chi2(P["data"], P["err"], orc.vis2data(P["npow"], P["nbis"], vis_o, P["uvpnt"], P["sign"],
                                                           P["phasor"]))
    m = _moves(img, W, 45, 5)
    d_o, s_o = _oracle_moves(orc, P, img64, vis_o, F_o, m)
    with GpairContext.from_problem(prob, image=img) as ctx:
        ctx.set_precision(prec)
        d, s = ctx.pixel_propose(m)
        d2, s2 = ctx.pixel_propose(m[:13])   # a second call, fewer moves, same buffers
    err = np.abs(d - d_o)
    worst = (err / (np.abs(d_o) + 1e-300)).max()
    print(f"  dchi2 worst rel {worst:.2e} (|dchi2| {np.abs(d_o).min():.3g} .. {np.abs(d_o).max():.3g}, "
          f"chi2 {chi2_o:.4g}); dS max rel {np.abs(s - s_o).max() / np.abs(s_o).max():.2e}", flush=True)
    assert np.all(err <= 1e-3 * np.abs(d_o) + 1e-7 * chi2_o)
    assert np.all(np.abs(s - s_o) <= 1e-12 * np.abs(s_o) + 1e-12)
    assert d[1] == 0.0 and s[1] == 0.0   # the null move
    assert np.array_equal(d2, d[:13]) and np.array_equal(s2, s[:13])


@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "bf16x3"])
def test_apply_updates_state_and_image(prec, orc):
    """moves applied in order (one pixel changed twice) give the state of the
    changed image: chi2, entropy and flux against the oracle, and the current
    image holds the new pixels"""
    prob = synth.make_problem("cfg1", seed=4, W=128, ntel=7, nch=6, nsnap=5)
    P = orc.unpack(prob)
    img = _image(prob, 9)
    with GpairContext.from_problem(prob, image=img) as ctx:
        ctx.set_precision(prec)
        c0, S0, F0 = ctx.pixel_state()
        m = _moves(img, prob.W, 9, 6)
        m = np.concatenate([m, m[3:4]])
        m["flux_a"][-1] *= 0.5   # pixel m[3].pix_a changes twice
        ctx.pixel_apply(m)
        c1, S1, F1 = ctx.pixel_state()
        new_img = img.copy()
        for mv in m:
            new_img[mv["pix_a"]] = mv["flux_a"]
            if mv["pix_b"] >= 0:
                new_img[mv["pix_b"]] = mv["flux_b"]
        assert np.array_equal(ctx.get("ci"), new_img)
        ctx.pixel_begin()   # the state of the new image rebuilt from scratch
        c2, _, _ = ctx.pixel_state()
    J, chi2_o, S_o, *_ = orc.objective(P, new_img.astype(np.float64), 0.0)
    F_o = new_img.astype(np.float64).sum()
    print(f"  state chi2 {abs(c1 - chi2_o) / chi2_o:.2e} (rebuilt {abs(c2 - chi2_o) / chi2_o:.2e}) "
          f"S {abs(S1 - S_o) / abs(S_o):.2e} F {abs(F1 - F_o) / F_o:.2e}", flush=True)
    assert abs(c1 - chi2_o) <= 2e-5 * chi2_o and abs(c2 - chi2_o) <= 2e-5 * chi2_o
    assert abs(S1 - S_o) <= 1e-9 * abs(S_o)
    assert abs(F1 - F_o) <= 1e-12 * F_o
    assert abs(F0 - img.astype(np.float64).sum()) <= 1e-12 * F0


def test_move_and_inverse_round_trip(orc):
    """full-size state (256 x 256, 36,864 uv): after applying a move, its inverse
    scores the opposite chi2 change; a null move scores 0; the tracked chi2 moves by
    the scored change"""
    prob = synth.make_problem("cfg2", seed=5)
    W = prob.W
    img = _image(prob, 10)
    with GpairContext.from_problem(prob, image=img) as ctx:
        ctx.set_precision(1)
        c0, _, _ = ctx.pixel_state()
        m = np.zeros(1, PIXEL_MOVE)
        m["pix_a"], m["pix_b"], m["flux_a"] = 100 + W * 120, 7 + W * 3, img[100 + W * 120] * 8.0
        m["flux_b"] = img[7 + W * 3] * 0.1
        d, s = ctx.pixel_propose(m)
        ctx.pixel_apply(m)
        c1, _, _ = ctx.pixel_state()
        inv = m.copy()
        inv["flux_a"], inv["flux_b"] = img[100 + W * 120], img[7 + W * 3]
        di, si = ctx.pixel_propose(inv)
        null = inv.copy()
        null["flux_a"], null["flux_b"] = m["flux_a"], m["flux_b"]
        dn, sn = ctx.pixel_propose(null)
    print(f"  move {d[0]:.6g} inverse {di[0]:.6g} tracked {c1 - c0:.6g}", flush=True)
    # the scored change of U is rounded to float (6e-8 of the change), the applied one is not
    assert abs(d[0] + di[0]) <= 1e-6 * abs(d[0]) + 1e-10 * c0
    assert abs(s[0] + si[0]) <= 1e-12 * abs(s[0])
    assert abs((c1 - c0) - d[0]) <= 1e-6 * abs(d[0]) + 1e-10 * c0
    assert dn[0] == 0.0 and sn[0] == 0.0


def test_state_follows_a_new_image(orc):
    """gpu_copy_image drops the pixel state: the next proposal scores moves against
    the new image"""
    prob = synth.make_problem("cfg1", seed=6)
    P = orc.unpack(prob)
    W = prob.W
    a, b = _image(prob, 11), _image(prob, 12)
    m = _moves(b, W, 8, 7)
    with GpairContext.from_problem(prob, image=a) as ctx:
        ctx.pixel_propose(m)
        ctx.gpu_copy_image(b)
        d, _ = ctx.pixel_propose(m)
    b64 = b.astype(np.float64)
    vis_o, F_o = orc.image2vis(W, P["au"], P["av"], b64)
    d_o, _ = _oracle_moves(orc, P, b64, vis_o, F_o, m)
    chi2_o = orc.chi2(P["data"], P["err"], orc.vis2data(P["npow"], P["nbis"], vis_o, P["uvpnt"], P["sign"],
                                                           P["phasor"]))
    assert np.all(np.abs(d - d_o) <= 1e-3 * np.abs(d_o) + 1e-7 * chi2_o)


def test_bad_moves_and_call_order_raise():
    prob = synth.make_problem("cfg1", seed=1)
    W = prob.W
    with GpairContext.from_problem(prob) as ctx:
        with pytest.raises(GpairError, match="before"):
            ctx.gpu_new_chi2()
        with pytest.raises(GpairError, match="outside"):
            ctx.gpu_update_vis_fluxchange(W, 0, 1.0)
        with pytest.raises(GpairError, match="same"):
            ctx.update_vis_positionchange(3, 3, 3, 3, 1.0)
        for bad in [(W * W, -1, 1.0, 0.0), (5, 5, 1.0, 1.0), (5, -2, 1.0, 0.0), (5, -1, float("nan"), 0.0)]:
            with pytest.raises(GpairError, match="invalid"):
                ctx.pixel_propose([bad])
        with pytest.raises(GpairError, match="moves"):
            ctx.pixel_propose(np.zeros(4097, PIXEL_MOVE))
        with pytest.raises(GpairError, match="moves"):
            ctx.pixel_apply(np.zeros(0, PIXEL_MOVE))
        # a rejected call changes nothing: a good call after it works
        d, _ = ctx.pixel_propose([(5, -1, float(prob.start[5]), 0.0)])
        assert d[0] == 0.0
