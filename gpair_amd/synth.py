"""Seeded synthetic interferometry problems in the reference's packed layout.

The reference reads OIFITS files (gpair1/trunk/getoifits.c:65-357) and packs
them with init_data (gpair1/trunk/gpair.c:997-1066).  This module builds the
same packed arrays directly from a seeded array layout and sky model:

* uv list: every V2 baseline first (``npow == nuv``), flipped into the +u half
  plane as getoifits.c:147-151 does; every T3 leg (ab, bc, ca) points at the V2
  baseline of the same snapshot/channel with ``sign = -1`` when the leg is the
  conjugate of the stored point (getoifits.c:218-337 semantics, no dedup
  search needed because legs are known by construction).
* data / err / phasor packed exactly like init_data:
  data = [V2 | (|T3amp|, 0) per T3], err = [1/|sigV2| | (1/sigamp, 1/|amp*sigphi*pi/180|)],
  phasor = exp(-i * t3phi * pi/180).

Noise conventions (recorded in ``Problem.meta``):
  sigma_V2  = max(0.02 * V2_true, 1e-4)      (2 % relative, absolute floor)
  sigma_amp = 0.05 * |T3_true|                (5 % relative)
  sigma_phi = 1 degree

Skies are uniform / limb-darkened disks, Gaussians, rings and points.  By
default the sky is rendered on the reconstruction grid and the data are the
FP64 DFT of that image (chi2/ndof of the truth ~ 1); closed-form visibilities of
the continuous sky are available with ``vis_model="analytic"``.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
from scipy import special

# gpair1/trunk/getoifits.h:87-88
PI_REF = 3.14159265358979323
RPMAS = (PI_REF / 180.0) / 3600000.0


@dataclasses.dataclass
class Problem:
    W: int
    pixellation: float            # mas per pixel
    u: np.ndarray                 # float32 [nuv] (wavelengths)
    v: np.ndarray                 # float32 [nuv]
    npow: int
    nbis: int
    data: np.ndarray              # float32 [npow + 2 nbis]
    err: np.ndarray               # float32 [npow + 2 nbis] (inverse errors)
    phasor: np.ndarray            # complex64 [nbis]
    uvpnt: np.ndarray             # int64 [nbis, 3]
    sign: np.ndarray              # int16 [nbis, 3]
    model: np.ndarray             # float32 [W*W] prior (default model)
    start: np.ndarray             # float32 [W*W] starting image
    truth_vis: np.ndarray         # complex128 [nuv] noiseless visibilities
    meta: dict
    truth: np.ndarray = None      # float32 [W*W] sky on the grid, scaled to the prior flux

    @property
    def nuv(self) -> int:
        return int(self.u.shape[0])

    @property
    def ndof(self) -> int:
        return self.npow + 2 * self.nbis

    @property
    def alpha_px(self) -> float:
        """radians per pixel, RPMAS * pixellation with the float32 pixellation."""
        return RPMAS * float(np.float32(self.pixellation))

    def au(self) -> np.ndarray:
        return self.alpha_px * self.u.astype(np.float64)

    def av(self) -> np.ndarray:
        return self.alpha_px * self.v.astype(np.float64)

    def uvpnt4(self) -> np.ndarray:
        """reference cl_long4 layout (gpair1/trunk/gpair.c:585-592)."""
        out = np.zeros((self.nbis, 4), np.int64)
        out[:, :3] = self.uvpnt
        return out

    def sign4(self) -> np.ndarray:
        """reference cl_short4 layout (gpair1/trunk/gpair.c:594-597)."""
        out = np.zeros((self.nbis, 4), np.int16)
        out[:, :3] = self.sign
        return out


# ---------------------------------------------------------------------------
# sky components: visibility of a unit-flux component centred at the origin
# ---------------------------------------------------------------------------
def _rho_rad(u, v):
    return np.hypot(u, v)  # wavelengths


def vis_uniform_disk(u, v, radius_mas):
    x = math.pi * 2.0 * radius_mas * RPMAS * _rho_rad(u, v)
    out = np.ones_like(x)
    nz = x > 1e-12
    out[nz] = 2.0 * special.j1(x[nz]) / x[nz]
    return out


def vis_ld_disk(u, v, radius_mas, mu):
    """linear limb darkening I(r) = 1 - mu (1 - sqrt(1 - r^2/R^2))."""
    x = math.pi * 2.0 * radius_mas * RPMAS * _rho_rad(u, v)
    out = np.ones_like(x)
    nz = x > 1e-6
    xs = x[nz]
    a = special.j1(xs) / xs
    b = (np.sin(xs) - xs * np.cos(xs)) / xs ** 3
    out[nz] = ((1.0 - mu) * a + mu * b) / (0.5 - mu / 6.0)
    return out


def vis_gauss(u, v, fwhm_mas):
    t = math.pi * fwhm_mas * RPMAS * _rho_rad(u, v)
    return np.exp(-t * t / (4.0 * math.log(2.0)))


def vis_ring(u, v, radius_mas):
    return special.j0(2.0 * math.pi * radius_mas * RPMAS * _rho_rad(u, v))


def shift(u, v, x_mas, y_mas):
    """phase of a component at (x, y) mas in the image frame of the DFT:
    X = exp(+2 pi i a u x), Y = exp(-2 pi i a v y) (gpair1/trunk/gpair.c:242-243)."""
    return np.exp(2j * math.pi * RPMAS * (u * x_mas - v * y_mas))


# ---------------------------------------------------------------------------
# array geometry
# ---------------------------------------------------------------------------
def _array_uv(rng, ntel, nch, nsnap, uvmax):
    """station uv coordinates s[snap, ch, tel, 2] (wavelengths) with every
    baseline inside +-uvmax on each axis."""
    pos = rng.uniform(-0.5, 0.5, size=(ntel, 2))
    pos -= pos.mean(axis=0)
    lam = np.linspace(1.0, 1.6, nch)              # relative wavelengths
    ang = np.linspace(0.0, math.pi / 2.0, nsnap, endpoint=False) + rng.uniform(0, 0.3)
    s = np.empty((nsnap, nch, ntel, 2))
    for k, a in enumerate(ang):
        rot = np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
        p = pos @ rot.T
        for c, l in enumerate(lam):
            s[k, c] = p / l
    # scale so that the longest baseline component reaches uvmax
    i, j = np.triu_indices(ntel, 1)
    b = s[:, :, j, :] - s[:, :, i, :]
    s *= uvmax / np.abs(b).max()
    return s


def build_layout(rng, ntel, nch, nsnap, uvmax):
    """V2 points and T3 references for all baselines and all triangles."""
    s = _array_uv(rng, ntel, nch, nsnap, uvmax)
    bi, bj = np.triu_indices(ntel, 1)
    nbase = bi.size
    base_index = {}
    for b in range(nbase):
        base_index[(bi[b], bj[b])] = b
    # V2 uv points ordered [snap][ch][baseline]
    uvl = s[:, :, bj, :] - s[:, :, bi, :]          # [snap, ch, nbase, 2]
    uvl = uvl.reshape(-1, 2)
    flip = uvl[:, 0] < 0.0                         # getoifits.c:147-151
    stored = np.where(flip[:, None], -uvl, uvl).astype(np.float32)
    tri = [(a, b, c) for a in range(ntel) for b in range(a + 1, ntel) for c in range(b + 1, ntel)]
    ntri = len(tri)
    uvpnt = np.empty((nsnap, nch, ntri, 3), np.int64)
    sign = np.empty((nsnap, nch, ntri, 3), np.int16)
    for k in range(nsnap):
        for c in range(nch):
            off = (k * nch + c) * nbase
            for t, (a, b, cc) in enumerate(tri):
                # legs ab = s_b - s_a, bc = s_c - s_b, ca = s_a - s_c
                for leg, (p, q) in enumerate(((a, b), (b, cc), (cc, a))):
                    lo, hi = min(p, q), max(p, q)
                    idx = off + base_index[(lo, hi)]
                    same_dir = (p, q) == (lo, hi)  # leg == stored baseline before the flip
                    sgn = 1 if same_dir else -1
                    if flip[idx]:
                        sgn = -sgn
                    uvpnt[k, c, t, leg] = idx
                    sign[k, c, t, leg] = sgn
    return stored, uvpnt.reshape(-1, 3), sign.reshape(-1, 3), (nbase, ntri)


# ---------------------------------------------------------------------------
# sky models for the configurations
# ---------------------------------------------------------------------------
def sky_components(kind, rng, fov_mas, pix):
    """list of component dicts (kind, flux, size parameters, x, y in mas from the
    image centre); fluxes sum to 1.  Point sources sit on pixel centres."""
    comps = []

    def snap(v):
        return round(v / pix) * pix

    if kind == "disk_spots":  # uniform disk r = 1.5 mas + 2 Gaussian spots (FWHM 0.3 mas)
        fs = rng.uniform(0.10, 0.30, size=2)
        comps.append(dict(kind="disk", flux=1.0 - fs.sum(), R=1.5, x=0.0, y=0.0))
        for f in fs:
            r, th = rng.uniform(0.2, 1.2), rng.uniform(0, 2 * math.pi)
            comps.append(dict(kind="gauss", flux=f, fwhm=0.3, x=r * math.cos(th), y=r * math.sin(th)))
    elif kind == "ld_disk_spots_companion":
        nsp = int(rng.integers(1, 6))
        fs = rng.uniform(0.03, 0.10, size=nsp)
        fc = rng.uniform(0.01, 0.05)
        R = 0.2 * fov_mas
        comps.append(dict(kind="ld_disk", flux=1.0 - fs.sum() - fc, R=R, mu=0.4, x=0.0, y=0.0))
        for f in fs:
            r, th = rng.uniform(0.1, 0.8) * R, rng.uniform(0, 2 * math.pi)
            comps.append(dict(kind="gauss", flux=f, fwhm=rng.uniform(0.15, 0.4) * R, x=r * math.cos(th),
                              y=r * math.sin(th)))
        r, th = rng.uniform(0.3, 0.42) * fov_mas, rng.uniform(0, 2 * math.pi)
        comps.append(dict(kind="point", flux=fc, x=snap(r * math.cos(th)), y=snap(r * math.sin(th))))
    elif kind == "learn_mix":
        # family for the learned start: a disk, limb-darkened disk, ring (with an
        # inner disk) or 2-4 Gaussian blobs, 0-3 Gaussian spots on disks, 0-3 point
        # sources in the outer field
        R = rng.uniform(0.12, 0.22) * fov_mas
        x0, y0 = rng.uniform(-0.03, 0.03, size=2) * fov_mas
        base = ("disk", "ld_disk", "ring", "blobs")[int(rng.integers(4))]
        nsp = int(rng.integers(0, 4)) if base in ("disk", "ld_disk") else 0
        npt = int(rng.integers(0, 4))
        fs = rng.uniform(0.03, 0.10, size=nsp)
        fp = rng.uniform(0.01, 0.05, size=npt)
        fb = 1.0 - fs.sum() - fp.sum()
        if base == "disk":
            comps.append(dict(kind="disk", flux=fb, R=R, x=x0, y=y0))
        elif base == "ld_disk":
            comps.append(dict(kind="ld_disk", flux=fb, R=R, mu=rng.uniform(0.2, 0.6), x=x0, y=y0))
        elif base == "ring":
            fr = rng.uniform(0.4, 0.8)
            comps.append(dict(kind="ring", flux=fb * fr, R=R, width=rng.uniform(0.04, 0.1) * R, x=x0, y=y0))
            comps.append(dict(kind="disk", flux=fb * (1 - fr), R=rng.uniform(0.2, 0.5) * R,
                              x=x0 + rng.uniform(-0.2, 0.2) * R, y=y0 + rng.uniform(-0.2, 0.2) * R))
        else:
            nb = int(rng.integers(2, 5))
            w = rng.uniform(0.5, 1.0, size=nb)
            for f in fb * w / w.sum():
                r, th = rng.uniform(0.0, 1.0) * R, rng.uniform(0, 2 * math.pi)
                comps.append(dict(kind="gauss", flux=f, fwhm=rng.uniform(0.3, 0.8) * R, x=x0 + r * math.cos(th),
                                  y=y0 + r * math.sin(th)))
        for f in fs:
            r, th = rng.uniform(0.1, 0.8) * R, rng.uniform(0, 2 * math.pi)
            comps.append(dict(kind="gauss", flux=f, fwhm=rng.uniform(0.15, 0.4) * R, x=x0 + r * math.cos(th),
                              y=y0 + r * math.sin(th)))
        for f in fp:
            r, th = rng.uniform(0.25, 0.42) * fov_mas, rng.uniform(0, 2 * math.pi)
            comps.append(dict(kind="point", flux=f, x=snap(r * math.cos(th)), y=snap(r * math.sin(th))))
    elif kind == "ring_disk":
        R = 0.2 * fov_mas
        fr = rng.uniform(0.3, 0.6)
        comps.append(dict(kind="ring", flux=fr, R=R, width=0.06 * R, x=0.0, y=0.0))
        comps.append(dict(kind="disk", flux=1.0 - fr, R=0.4 * R, x=0.1 * R, y=-0.05 * R))
    else:
        raise ValueError(kind)
    return comps


def component_vis(c, u, v):
    """analytic visibility of one component (flux-weighted, position phase included)."""
    k = c["kind"]
    if k == "disk":
        base = vis_uniform_disk(u, v, c["R"])
    elif k == "ld_disk":
        base = vis_ld_disk(u, v, c["R"], c["mu"])
    elif k == "gauss":
        base = vis_gauss(u, v, c["fwhm"])
    elif k == "point":
        base = np.ones_like(u)
    elif k == "ring":
        base = vis_ring(u, v, c["R"]) * vis_gauss(u, v, c["width"])
    else:
        raise ValueError(k)
    return c["flux"] * base * shift(u, v, c["x"], c["y"])


def render(comps, W, pix, oversample=4):
    """sky image on the W x W grid, image[x + W*y], pixel (x, y) at ((x - W/2) pix,
    (y - W/2) pix) mas; each extended component is box-averaged over
    oversample^2 sub-pixels and normalised to its flux."""
    o = oversample
    sub = (np.arange(W * o) + 0.5) / o - 0.5 - W // 2      # sub-pixel centres in pixels
    X, Y = np.meshgrid(sub * pix, sub * pix)               # [y, x] in mas
    img = np.zeros((W, W))
    for c in comps:
        if c["kind"] == "point":
            ix = int(round(c["x"] / pix)) + W // 2
            iy = int(round(c["y"] / pix)) + W // 2
            img[iy, ix] += c["flux"]
            continue
        dx, dy = X - c["x"], Y - c["y"]
        r2 = dx * dx + dy * dy
        if c["kind"] == "disk":
            f = (r2 <= c["R"] ** 2).astype(np.float64)
        elif c["kind"] == "ld_disk":
            t = np.clip(1.0 - r2 / c["R"] ** 2, 0.0, None)
            f = np.where(r2 <= c["R"] ** 2, 1.0 - c["mu"] * (1.0 - np.sqrt(t)), 0.0)
        elif c["kind"] == "gauss":
            f = np.exp(-4.0 * math.log(2.0) * r2 / c["fwhm"] ** 2)
        elif c["kind"] == "ring":
            s = c["width"] / (2.0 * math.sqrt(2.0 * math.log(2.0)))
            f = np.exp(-(np.sqrt(r2) - c["R"]) ** 2 / (2.0 * s * s))
        else:
            raise ValueError(c["kind"])
        f = f.reshape(W, o, W, o).mean(axis=(1, 3))
        img += c["flux"] * f / f.sum()
    return img.ravel()


def dft64(img, W, au, av, chunk=4096):
    """FP64 separable DFT V_p = (1/F) sum_y Y[p,y] sum_x X[p,x] I[y][x] of a grid image."""
    I2 = img.reshape(W, W)
    F = img.sum()
    xs = np.arange(W, dtype=np.float64)
    out = np.empty(au.size, np.complex128)
    for a in range(0, au.size, chunk):
        b = min(a + chunk, au.size)
        tx = np.outer(au[a:b], xs)
        ty = np.outer(-av[a:b], xs)
        Xp = np.exp(2j * math.pi * (tx - np.rint(tx)))
        Yp = np.exp(2j * math.pi * (ty - np.rint(ty)))
        T = Xp @ I2.T                     # T[p, y] = sum_x X[p, x] I[y, x]
        out[a:b] = (T * Yp).sum(axis=1) / F
    return out


CONFIGS = {
    # reference CPU-runnable case: 64x64 @ 0.1 mas, 6 telescopes, 8 ch x 4 snapshots
    "cfg1": dict(W=64, pix=0.1, ntel=6, nch=8, nsnap=4, uvmax=3e8, sky="disk_spots", alpha=0.025),
    # headline: 256x256 @ 0.05 mas, 9 telescopes, 64 ch x 16 snapshots
    # prior: the reference's default model, a Gaussian of FWHM 5 mas (gpair1/trunk/gpair.c:82-84), centred
    "cfg2": dict(W=256, pix=0.05, ntel=9, nch=64, nsnap=16, uvmax=3e8, sky="ld_disk_spots_companion", alpha=0.3,
                 prior=("gauss", 5.0)),
    # precision variant: 512x512, 256 channels
    # learned start: skies from the learn_mix family on the headline grid (256x256 @
    # 0.05 mas) sampled by the 6-telescope array of cfg1 (480 V2, 640 closure triangles
    # before the cut below); alpha 0.03 for this sparse sampling (0.3 leaves chi2/ndof
    # near 2 on some skies); closure triangles with |T3| < 1e-3 are not observed: smooth
    # skies have closure amplitudes far below any instrument's reach, and the reference
    # weights the closure phase by 1 / |observed amplitude|
    "cfg5": dict(W=256, pix=0.05, ntel=6, nch=8, nsnap=4, uvmax=3e8, sky="learn_mix", alpha=0.03,
                 prior=("gauss", 5.0), t3_min=1e-3),
    # batching: 9 telescopes (T3 = 84/36 of V2), widths and V2 counts per problem
    # drawn by batch_problems
    "cfg4": dict(W=256, pix=0.05, ntel=9, nch=16, nsnap=4, uvmax=3e8, sky="learn_mix", alpha=0.3,
                 prior=("gauss", 5.0), t3_min=1e-3),
    "cfg3": dict(W=512, pix=0.05, ntel=9, nch=256, nsnap=16, uvmax=3e8, sky="ring_disk", alpha=0.3,
                 prior=("gauss", 10.0)),
}


def set_model(W, pix, mtype, width, flux, centered=True):
    """prior image through the library's gpair_set_model (gpair1/trunk/cpu.c:645-749)."""
    import ctypes as C
    from . import _lib
    out = np.zeros(W * W, np.float32)
    rc = _lib.load().gpair_set_model(out.ctypes.data, int(W), C.c_float(pix), int(mtype), C.c_float(width),
                                     C.c_float(flux), int(bool(centered)))
    if rc != 0:
        raise ValueError(f"gpair_set_model: bad arguments (type {mtype})")
    return out


def make_problem(cfg="cfg1", seed=1, W=None, pix=None, ntel=None, nch=None, nsnap=None, uvmax=None,
                 sky=None, prior_flux=None, noise=True, vis_model="grid", prior=None, dft=None) -> Problem:
    """Seeded synthetic problem.  Named configs fix the sizes; keyword
    overrides allow small test shapes with the same construction.

    vis_model: "grid"     -- the sky is rendered on the reconstruction grid and its
                             visibilities are the FP64 DFT of that image, so the
                             truth image reaches chi2/ndof ~ 1 (default);
               "analytic" -- closed-form visibilities of the continuous sky.
    prior_flux: total flux of the flat prior (default W*W: mean prior pixel 1).
    dft: replacement for dft64 with the same signature (e.g. an FP64 device DFT).
    meta["alpha"] is the entropy weight used with this prior scale: J(cI; cM)
    with alpha/c equals J(I; M) with alpha, so alpha scales inversely with the
    prior flux (steps of the line search scale with its square)."""
    c = dict(CONFIGS[cfg])
    for k, val in (("W", W), ("pix", pix), ("ntel", ntel), ("nch", nch), ("nsnap", nsnap),
                   ("uvmax", uvmax), ("sky", sky)):
        if val is not None:
            c[k] = val
    rng = np.random.default_rng(seed)
    uv, uvpnt, sign, (nbase, ntri) = build_layout(rng, c["ntel"], c["nch"], c["nsnap"], c["uvmax"])
    u = uv[:, 0].astype(np.float64)
    v = uv[:, 1].astype(np.float64)
    W_ = c["W"]
    pixf = float(np.float32(c["pix"]))
    fov = W_ * pixf
    comps = sky_components(c["sky"], rng, fov, pixf)
    truth = render(comps, W_, pixf)
    alpha_px = RPMAS * pixf
    if vis_model == "grid":
        V = (dft or dft64)(truth, W_, alpha_px * u, alpha_px * v)
    elif vis_model == "analytic":
        xc = (W_ // 2) * pixf     # image centre in the DFT frame
        V = np.zeros(u.shape, np.complex128)
        for comp in comps:
            V += component_vis(dict(comp, x=comp["x"] + xc, y=comp["y"] + xc), u, v)
    else:
        raise ValueError(vis_model)
    npow = u.size
    nbis = uvpnt.shape[0]
    v2 = np.abs(V) ** 2
    legs = np.where(sign < 0, np.conj(V[uvpnt]), V[uvpnt])
    T3 = legs.prod(axis=1)
    if c.get("t3_min", 0.0) > 0.0:
        keep = np.abs(T3) >= c["t3_min"]
        uvpnt, sign, T3 = uvpnt[keep], sign[keep], T3[keep]
        nbis = uvpnt.shape[0]
    sig_v2 = np.maximum(0.02 * v2, 1e-4)
    sig_amp = np.maximum(0.05 * np.abs(T3), 1e-30)
    sig_phi = np.full(nbis, 1.0)
    if noise:
        v2_obs = v2 + rng.normal(size=npow) * sig_v2
        amp_obs = np.abs(T3) + rng.normal(size=nbis) * sig_amp
        phi_obs = np.degrees(np.angle(T3)) + rng.normal(size=nbis) * sig_phi
    else:
        v2_obs, amp_obs, phi_obs = v2, np.abs(T3), np.degrees(np.angle(T3))
    v2_obs = v2_obs.astype(np.float32)
    amp_obs = amp_obs.astype(np.float32)
    phi_obs = phi_obs.astype(np.float32)
    sig_v2 = sig_v2.astype(np.float32)
    sig_amp = sig_amp.astype(np.float32)
    sig_phi = sig_phi.astype(np.float32)
    # init_data packing, gpair1/trunk/gpair.c:1000-1063 (float arithmetic)
    ndof = npow + 2 * nbis
    data = np.zeros(ndof, np.float32)
    err = np.zeros(ndof, np.float32)
    data[:npow] = v2_obs
    err[:npow] = 1.0 / np.abs(sig_v2)
    data[npow::2] = np.abs(amp_obs)
    err[npow::2] = 1.0 / sig_amp
    err[npow + 1::2] = 1.0 / np.abs(amp_obs.astype(np.float64) * sig_phi * PI_REF / 180.0)
    phasor = np.exp(-1j * phi_obs.astype(np.float64) * PI_REF / 180.0).astype(np.complex64)
    # flat prior of the given total flux (set_model type 0, gpair1/trunk/cpu.c:653-662),
    # start image = max(prior, 1e-8) (gpair1/trunk/gpair.c:217-225)
    if prior_flux is None:
        prior_flux = float(W_ * W_)
    if prior is None:
        prior = c.get("prior", ("flat", 0.0))
    if prior[0] == "flat":
        model = np.full(W_ * W_, prior_flux / float(W_ * W_), np.float32)
    else:
        model = set_model(W_, pixf, {"dirac": 1, "disk": 2, "gauss": 3, "lorentz": 4}[prior[0]], prior[1],
                          prior_flux, centered=True)
    start = np.maximum(model, np.float32(1e-8))
    meta = dict(cfg=cfg, seed=seed, nbase=nbase, ntri=ntri, sky=c["sky"], W=W_, pix=c["pix"],
                noise="sigV2=max(0.02*V2,1e-4); sigamp=0.05*|T3|; sigphi=1deg",
                t3_min=c.get("t3_min", 0.0),
                uvmax=c["uvmax"], ntel=c["ntel"], nch=c["nch"], nsnap=c["nsnap"], vis_model=vis_model,
                prior_flux=prior_flux, components=comps, alpha=c["alpha"], prior=tuple(prior))
    return Problem(W=W_, pixellation=c["pix"], u=uv[:, 0].copy(), v=uv[:, 1].copy(), npow=npow,
                   nbis=nbis, data=data, err=err, phasor=phasor, uvpnt=uvpnt, sign=sign,
                   model=model, start=start, truth_vis=V, meta=meta, truth=(truth * prior_flux).astype(np.float32))


BATCH_WIDTHS = (64, 128, 192, 256)


def batch_spec(count=64, seed=4, v2_range=(500, 40000)):
    """sizes of the batching workload: per problem a width from BATCH_WIDTHS and a
    V2 count near a uniform draw from v2_range (36 baselines x channels x snapshots)."""
    rng = np.random.default_rng(seed)
    spec = []
    for i in range(count):
        W = int(rng.choice(BATCH_WIDTHS))
        target = rng.uniform(*v2_range)
        nsnap = int(rng.integers(1, 17))
        nch = max(1, int(round(target / (36 * nsnap))))
        spec.append(dict(seed=seed * 1000 + i, W=W, nch=nch, nsnap=nsnap))
    return spec


def batch_problems(count=64, seed=4, v2_range=(500, 40000), workers=8):
    """cfg4 problems (built on host threads: the FP64 numpy DFTs release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    spec = batch_spec(count, seed, v2_range)
    with ThreadPoolExecutor(max_workers=workers) as pool:
        return list(pool.map(lambda k: make_problem("cfg4", **k), spec))
