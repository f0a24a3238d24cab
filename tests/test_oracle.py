"""Pins the FP64 oracle before it is trusted as the checker.

* golden vectors written by the reference's own compiled cpu.c (tests/golden),
  and the live reference library when oracle/_ref is built;
* closed forms (point sources), brute-force DFT on a tiny grid, translation
  invariance, the scale-invariance identity sum(I * grad chi2) = 0, the adjoint
  dot-product identity, central finite differences, entropy identities;
* the separable adjoint against the reference's per-triangle formula.
A planted mistake (sign flip in the gradient) is shown to fail the FD check.
"""
import glob
import math
import os

import numpy as np
import pytest

from gpair_amd import synth
from oracle import Oracle, RefCPU, ref_available

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_cpu_*.npz")))


@pytest.fixture(scope="module")
def orc():
    return Oracle()


class _Fixture:
    """a golden file presented like a synth.Problem"""

    def __init__(self, path):
        z = np.load(path)
        self.z = z
        self.W = int(z["W"])
        self.pixellation = float(z["pix"])
        self.u, self.v = z["u"], z["v"]
        self.npow, self.nbis = int(z["npow"]), int(z["nbis"])
        self.data, self.err, self.phasor = z["data"], z["err"], z["phasor"]
        self.uvpnt, self.sign, self.model = z["uvpnt"], z["sign"], z["model"]

    @property
    def nuv(self):
        return self.u.size

    def au(self):
        return synth.RPMAS * float(np.float32(self.pixellation)) * self.u.astype(np.float64)

    def av(self):
        return synth.RPMAS * float(np.float32(self.pixellation)) * self.v.astype(np.float64)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_oracle_matches_reference_golden_vectors(path, orc):
    """FP64 oracle vs the reference's float outputs (cpu.c) on the same inputs.
    Bounds are float-precision: the reference accumulates in float."""
    f = _Fixture(path)
    z = f.z
    P = orc.unpack(f)
    img = z["image"].astype(np.float64)
    J, chi2, S, vis, mock, F = orc.objective(P, img, 1.0)
    assert abs(chi2 - float(z["ref_chi2"])) / chi2 < 2e-5
    assert np.abs(vis - z["ref_vis"]).max() < 5e-6 * np.abs(vis).max()
    assert np.abs(mock - z["ref_mock"]).max() < 5e-6 * np.abs(mock).max()
    g = orc.grad_direct(P, img, vis, mock)
    assert np.abs(g - z["ref_grad"]).max() < 5e-5 * np.abs(g).max()
    assert abs(S - float(z["ref_entropy"])) / abs(S) < 1e-5
    assert np.allclose(orc.entropy_grad(img, P["model"]), z["ref_entropy_grad"], rtol=1e-5, atol=1e-6)


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built")
def test_oracle_matches_live_reference(orc):
    prob = synth.make_problem("cfg1", seed=21, ntel=5, nch=4, nsnap=3)
    R = RefCPU(prob)
    P = orc.unpack(prob)
    img = (prob.start * (1 + np.random.default_rng(3).random(prob.W ** 2))).astype(np.float32)
    chi2_r, vis_r, mock_r = R.image2chi2(img)
    g_r = R.data_gradient(img)
    J, chi2, S, vis, mock, F = orc.objective(P, img.astype(np.float64), 1.0)
    assert abs(chi2 - chi2_r) / chi2 < 2e-5
    assert np.abs(vis - vis_r).max() < 5e-6 * np.abs(vis).max()
    g = orc.grad_direct(P, img.astype(np.float64), vis, mock)
    assert np.abs(g - g_r).max() < 5e-5 * np.abs(g).max()


def _small(seed=1, **kw):
    kw.setdefault("ntel", 5)
    kw.setdefault("nch", 3)
    kw.setdefault("nsnap", 2)
    return synth.make_problem("cfg1", seed=seed, **kw)


def test_point_source_closed_form(orc):
    prob = _small(2)
    P = orc.unpack(prob)
    W = prob.W
    x0, y0 = 17, 40
    img = np.zeros(W * W)
    img[x0 + W * y0] = 3.0
    vis, F = orc.image2vis(W, P["au"], P["av"], img)
    expect = np.exp(2j * math.pi * (P["au"] * x0 - P["av"] * y0))
    assert np.abs(vis - expect).max() < 1e-12
    mock = orc.vis2data(P["npow"], P["nbis"], vis, P["uvpnt"], P["sign"], P["phasor"])
    assert np.allclose(mock[:P["npow"]], 1.0, atol=1e-12)
    # closure: the three legs of a point source multiply to 1, so T3 = phasor
    T = mock[P["npow"]::2] + 1j * mock[P["npow"] + 1::2]
    assert np.abs(T - prob.phasor.astype(np.complex128)).max() < 1e-4   # float32 uv closure


def test_two_points_and_brute_force_dft(orc):
    prob = _small(3)
    P = orc.unpack(prob)
    W = prob.W
    rng = np.random.default_rng(1)
    img = rng.random(W * W)
    vis, F = orc.image2vis(W, P["au"], P["av"], img)
    x = np.arange(W)
    # brute force over all pixels for 8 uv points
    for p in range(0, prob.nuv, max(1, prob.nuv // 8)):
        E = np.exp(2j * math.pi * (P["au"][p] * x[None, :] - P["av"][p] * x[:, None]))
        assert abs((img.reshape(W, W) * E).sum() / img.sum() - vis[p]) < 1e-12
    img2 = np.zeros(W * W)
    img2[5 + W * 9] = 1.0
    img2[30 + W * 2] = 2.0
    vis2, _ = orc.image2vis(W, P["au"], P["av"], img2)
    e1 = np.exp(2j * math.pi * (P["au"] * 5 - P["av"] * 9))
    e2 = np.exp(2j * math.pi * (P["au"] * 30 - P["av"] * 2))
    assert np.abs(vis2 - (e1 + 2 * e2) / 3).max() < 1e-12


def test_translation_invariance(orc):
    prob = _small(4)
    P = orc.unpack(prob)
    W = prob.W
    rng = np.random.default_rng(2)
    base = np.zeros((W, W))
    base[20:40, 18:36] = rng.random((20, 18))
    shifted = np.roll(np.roll(base, 7, axis=0), -5, axis=1)
    m = []
    for im in (base, shifted):
        vis, _ = orc.image2vis(W, P["au"], P["av"], im.ravel())
        m.append(orc.vis2data(P["npow"], P["nbis"], vis, P["uvpnt"], P["sign"], P["phasor"]))
    npow = P["npow"]
    assert np.abs(m[0][:npow] - m[1][:npow]).max() < 1e-12
    # T3 legs close only to float32 rounding of the stored u, v
    assert np.abs(m[0][npow:] - m[1][npow:]).max() < 1e-6 * np.abs(m[0][npow:]).max()


def test_gradient_identities_and_fd(orc):
    prob = _small(5)
    P = orc.unpack(prob)
    W = prob.W
    rng = np.random.default_rng(4)
    img = prob.start.astype(np.float64) * (0.5 + rng.random(W * W))
    J, chi2, S, vis, mock, F = orc.objective(P, img, 0.0)
    coef, c0 = orc.grad_coeffs(P, vis, mock)
    g_sep = orc.grad_separable(P, coef, c0, F)
    g_dir = orc.grad_direct(P, img, vis, mock)
    assert np.abs(g_sep - g_dir).max() < 1e-10 * np.abs(g_dir).max()
    # chi2 is invariant under I -> cI
    assert abs((img * g_sep).sum()) < 1e-9 * np.abs(img * g_sep).sum()
    # central finite differences along random directions

    def chi2_of(x):
        return orc.objective(P, x, 0.0)[1]

    for k in range(3):
        d = rng.standard_normal(W * W) * img
        h = 1e-5
        fd = (chi2_of(img + h * d) - chi2_of(img - h * d)) / (2 * h)
        an = g_sep @ d
        assert abs(fd - an) < 1e-6 * abs(an) + 1e-9 * chi2
    # a planted sign mistake is caught by the same check
    d = rng.standard_normal(W * W) * img
    fd = (chi2_of(img + 1e-5 * d) - chi2_of(img - 1e-5 * d)) / 2e-5
    assert abs(fd - (-g_sep) @ d) > 1e-3 * abs(fd)


def test_adjoint_dot_product(orc):
    """Re sum_p a_p (F V)_p == sum I * G(a) for the unnormalised DFT pair."""
    prob = _small(6)
    P = orc.unpack(prob)
    W = prob.W
    rng = np.random.default_rng(5)
    img = rng.random(W * W)
    a = rng.standard_normal(prob.nuv) + 1j * rng.standard_normal(prob.nuv)
    vis, F = orc.image2vis(W, P["au"], P["av"], img)
    lhs = np.real((a * vis * F).sum())
    G = orc.grad_separable(P, a, 0.0, 1.0)
    rhs = (img * G).sum()
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_chi2_and_entropy_identities(orc):
    prob = _small(7)
    P = orc.unpack(prob)
    assert orc.chi2(P["data"], P["err"], P["data"].copy()) == 0.0
    M = P["model"]
    assert abs(orc.entropy(M, M)) < 1e-12 * M.sum()
    assert np.abs(orc.entropy_grad(M, M)).max() < 1e-12
    rng = np.random.default_rng(6)
    img = M * (0.5 + rng.random(M.size))
    d = rng.standard_normal(M.size) * M
    h = 1e-6
    fd = (orc.entropy(img + h * d, M) - orc.entropy(img - h * d, M)) / (2 * h)
    assert abs(fd - orc.entropy_grad(img, M) @ d) < 1e-6 * abs(fd)
    # concave: S below its tangent
    s0, g0 = orc.entropy(img, M), orc.entropy_grad(img, M)
    assert orc.entropy(img + 0.3 * d, M) <= s0 + 0.3 * (g0 @ d)
    # stale value kept where the image is not positive (kernel_entropy_gs_grad.cl:18-23)
    img2 = img.copy()
    img2[:10] = 0.0
    old = np.full(M.size, 7.0)
    g2 = orc.entropy_grad(img2, M, old=old)
    assert np.all(g2[:10] == 7.0) and np.all(g2[10:] != 7.0)


def test_cg_oracle_runs_reference_rules(orc):
    """J never increases across accepted steps with PR (Armijo with d.g < 0),
    and a Python replay of gpair.c:682-915 / gpu.c:1947-2065 gives the same trace."""
    prob = _small(8)
    P = orc.unpack(prob)
    alpha = 0.03
    img, tr = orc.cg(P, prob.start.astype(np.float64), alpha, method=2, max_iter=6, chi2r_target=0.0)
    assert np.all(np.diff(tr[:, 0]) <= 1e-9 * np.abs(tr[:-1, 0]))
    rep = _replay(orc, P, prob.start.astype(np.float64), alpha, 6)
    assert np.allclose(rep, tr[:, 0], rtol=1e-9, atol=0)


def _replay(orc, P, img, alpha, nit):
    """independent Python restatement of the reference control flow (FP64)."""
    c1, c2, minv = 1e-4, 0.1, 1e-8
    img = img.copy()
    n = img.size
    Js = []
    g_old = np.zeros(n)
    d = np.zeros(n)

    def J_of(x):
        r = orc.objective(P, x, alpha)
        return r[0], r[3], r[4]

    def grad_of(x, vis, mock):
        return orc.gradient(P, x, alpha, vis, mock)[0]

    def step(s):
        return np.maximum(img + s * d, minv)

    def zoom(lo, hi, J_lo, w1, J0):
        s = s_old = J_old = 0.0
        cnt = 0
        while True:
            if cnt > 0 and (J_old - J0 - w1 * s_old) != 0.0:
                s = abs(w1 * s_old ** 2 / (2.0 * (J_old - J0 - w1 * s_old)))
            if cnt == 0 or s < lo or s > hi:
                s = (hi - lo) / 2.0 + lo
            if abs(hi - lo) < 1e-14:
                return lo
            t = step(s)
            Jt, vis, mock = J_of(t)
            if Jt > J0 + c1 * s * w1 or Jt >= J_lo:
                hi = s
            else:
                w2 = d @ grad_of(t, vis, mock)
                if w2 >= c2 * w1 or cnt > 10:
                    return s
                if w2 * (hi - lo) >= 0:
                    hi = lo
                lo = s
            s_old, J_old = s, Jt
            cnt += 1

    for it in range(nit):
        J, vis, mock = J_of(img)
        Js.append(J)
        g = grad_of(img, vis, mock)
        if it == 0:
            beta = 0.0
        else:
            gg, go, oo = g @ g, g @ g_old, g_old @ g_old
            beta = max(0.0, (gg - go) / oo)
            if abs(go / gg) > 0.5:
                beta = 0.0
        d = beta * d - g
        w1 = d @ g
        s, s_old, J0, J_old, lsit = 1.0, 0.0, J, J, 1
        while True:
            t = step(s)
            Jt, vis, mock = J_of(t)
            if Jt > J0 + c1 * s * w1 or (Jt >= J_old and lsit > 1):
                sel = zoom(s_old, s, J_old, w1, J0)
                break
            w2 = d @ grad_of(t, vis, mock)
            if abs(w2) <= -c2 * w1:
                sel = s
                break
            if w2 >= 0:
                sel = zoom(s, s_old, Jt, w1, J0)
                break
            s_old, J_old = s, Jt
            s = min(10 * s, 100.0)
            lsit += 1
        img = np.maximum(img + sel * d, minv)
        g_old = g
    return np.array(Js)


def test_parabola_step_closed_form(orc):
    """orc_parabola_step: the vertex of the parabola through three points."""
    a, b, c = 2.0, -3.0, 0.5          # J(s) = a (s - 0.75)^2 + c, vertex 0.75
    f = lambda s: a * (s - 0.75) ** 2 + c
    for s in ((0.1, 0.5, 1.3), (0.5, 0.7, 2.0), (0.0, 1.0, 4.0)):
        v = orc.L.orc_parabola_step(s[0], s[1], s[2], f(s[0]), f(s[1]), f(s[2]))
        assert abs(v - 0.75) <= 1e-12
    del b


def test_dft_raw_is_the_unnormalised_dft(orc):
    """orc_dft_raw = F * V: point source of flux 3 at (x0, y0) gives 3 X[p,x0] Y[p,y0]."""
    prob = _small(9)
    W = prob.W
    img = np.zeros(W * W)
    x0, y0 = 5, 11
    img[x0 + W * y0] = 3.0
    U = orc.dft_raw(W, prob.au(), prob.av(), img)
    ref = 3.0 * np.exp(2j * np.pi * (prob.au() * x0 - prob.av() * y0))
    assert np.abs(U - ref).max() <= 1e-12


def test_pcg_oracle_descends_and_tracks_the_objective(orc):
    """orc_pcg (the accelerated rules): its FP64 objective never increases, the
    objective it tracks by exact per-datum / per-pixel differences equals the
    objective recomputed from the final image, and the batch counters add up."""
    prob = _small(10)
    P = orc.unpack(prob)
    alpha = prob.meta["alpha"]
    img, tr = orc.pcg(P, prob.start.astype(np.float64), alpha, precond=1, max_iter=15, chi2r_target=0.0)
    assert len(tr) >= 10
    J = tr[:, 0]
    assert np.all(np.diff(J) <= 1e-12 * np.abs(J[:-1]))
    assert np.all(np.diff(tr[:, 5]) >= 1) and np.all(np.diff(tr[:, 6]) == 1)   # DFT counters
    # the objective tracked at the start of iteration 6 equals the objective of the
    # image after 6 iterations, recomputed from scratch
    img6, _ = orc.pcg(P, prob.start.astype(np.float64), alpha, precond=1, max_iter=6, chi2r_target=0.0)
    J6 = orc.objective(P, img6, alpha)[0]
    assert abs(tr[6, 0] - J6) <= 1e-10 * abs(J6)
    assert abs(orc.objective(P, img, alpha)[0] - J[-1]) > 0        # the final commit moved it
    # chi2 and entropy columns match the objective's parts at the first row
    J0, chi20, S0, *_ = orc.objective(P, prob.start.astype(np.float64), alpha)
    assert abs(tr[0, 1] - chi20) <= 1e-12 * chi20 and abs(tr[0, 2] - S0) <= 1e-9 * max(abs(S0), 1.0)


def test_pixel_update_is_the_dft_of_the_changed_image(orc):
    """orc_px_update (update_vis_fluxchange / update_vis_positionchange,
    gpair1/trunk/cpu.c:477-497): the visibilities after a one- or two-pixel change
    equal the DFT of the changed image, the new flux is the changed image's flux,
    and a one-pixel change has the reference's form r V + (1 - r) X Y with r the
    ratio of TOTAL fluxes.  The reference kernel's pixel-flux ratio and its y
    index taken from x_pos (kernel_vis_update_fluxchange.cl:27-28) both fail."""
    prob = _small(11)
    P = orc.unpack(prob)
    W = prob.W
    rng = np.random.default_rng(11)
    img = rng.random(W * W) + 0.05
    vis, F = orc.image2vis(W, P["au"], P["av"], img)
    x, y, new = 13, 29, 2.5
    pix = x + W * y
    d = new - img[pix]
    changed = img.copy()
    changed[pix] = new
    ref, Fref = orc.image2vis(W, P["au"], P["av"], changed)
    got, Fn = orc.px_update(P, vis, F, pix, d)
    assert np.abs(got - ref).max() < 1e-12 and abs(Fn - Fref) < 1e-12 * Fref
    XY = np.exp(2j * math.pi * (P["au"] * x - P["av"] * y))
    r = F / (F + d)
    assert np.abs(r * vis + (1 - r) * XY - ref).max() < 1e-12
    # planted mistakes: the ratio of pixel fluxes, the y table indexed by x
    r_pix = img[pix] / new
    assert np.abs(r_pix * vis + (1 - r_pix) * XY - ref).max() > 1e-3
    XX = np.exp(2j * math.pi * (P["au"] * x - P["av"] * x))
    assert np.abs(r * vis + (1 - r) * XX - ref).max() > 1e-3
    # position change: flux f from (3, 4) to (50, 7), total flux unchanged
    f, po, pn = 0.3, 3 + W * 4, 50 + W * 7
    moved = img.copy()
    moved[po] -= f
    moved[pn] += f
    ref2, _ = orc.image2vis(W, P["au"], P["av"], moved)
    got2, F2 = orc.px_update(P, vis, F, [po, pn], [-f, f])
    assert np.abs(got2 - ref2).max() < 1e-12 and abs(F2 - F) < 1e-12 * F
    e_old = np.exp(2j * math.pi * (P["au"] * 3 - P["av"] * 4))
    e_new = np.exp(2j * math.pi * (P["au"] * 50 - P["av"] * 7))
    assert np.abs(vis + f / F * (e_new - e_old) - ref2).max() < 1e-12


def test_pixel_propose_scores_moves_by_chi2_and_entropy_differences(orc):
    """orc_px_propose: per move, chi2 and entropy of the changed image minus those
    of the image, each recomputed from scratch (full DFT); a move that keeps the
    pixel value scores exactly zero."""
    prob = _small(12)
    P = orc.unpack(prob)
    W = prob.W
    rng = np.random.default_rng(12)
    img = prob.start.astype(np.float64) * (0.5 + rng.random(W * W))
    vis, F = orc.image2vis(W, P["au"], P["av"], img)
    K = 10
    pa = rng.choice(W * W, K, replace=False).astype(np.int32)
    pb = np.where(np.arange(K) % 3 == 0, (pa + 1 + rng.integers(0, W * W - 1, K)) % (W * W), -1).astype(np.int32)
    fa = img[pa] * (1 + rng.normal(0, 2, K)) ** 2
    fb = np.where(pb >= 0, img[np.maximum(pb, 0)] * 3.0, 0.0)
    fa[1] = img[pa[1]]
    pb[1] = -1
    dchi2, dS = orc.px_propose(P, img, vis, F, pa, pb, fa, fb)
    _, chi0, S0, *_ = orc.objective(P, img, 0.0)
    for k in range(K):
        new = img.copy()
        new[pa[k]] = fa[k]
        if pb[k] >= 0:
            new[pb[k]] = fb[k]
        _, chi, S, *_ = orc.objective(P, new, 0.0)
        assert abs(dchi2[k] - (chi - chi0)) <= 1e-9 * chi0
        assert abs(dS[k] - (S - S0)) <= 1e-9 * max(abs(S0), 1.0)
    assert dchi2[1] == 0.0 and dS[1] == 0.0
