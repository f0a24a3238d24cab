"""ctypes front end of the checker libraries.  TEST INFRASTRUCTURE ONLY.

* ``Oracle``  -- the FP64 restatement in gpair_oracle.c.
* ``RefCPU``  -- the reference's own CPU model (gpair1/trunk/cpu.c) compiled by
  ``make -C oracle ref`` into oracle/_ref/libgpair_refcpu.so.  Optional: absent
  when the reference tree was not available at build time.

Only tests, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libgpair_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libgpair_refcpu.so")
REF_SRC = "/root/reference/gpair1/trunk/cpu.c"

_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
# phi(user, s) callbacks of the line search (orc_linesearch / orc_ls_zoom)
PHI = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)

# exit codes of the reference line search (info[0], gpair_oracle.c)
LS_BRACKET, LS_ZOOM_CURV, LS_ZOOM_COUNT, LS_ZOOM_WIDTH, LS_CAP = 1, 2, 3, 4, 5
LS_INFO = ("exit", "s", "J", "w2", "nJ", "nG", "counter", "width")
_ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


def build(ref: bool | None = None) -> None:
    """compile the oracle (and, when the reference tree exists, oracle/_ref)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all"], check=True)
    if ref is None:
        ref = os.path.exists(REF_SRC)
    if ref:
        subprocess.run(["make", "-s", "-B", "-C", HERE, "ref"], check=True)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Oracle:
    """FP64 restatement; every method mirrors a function of gpair_oracle.c."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        self.L = L
        L.orc_tables.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_flux.argtypes = [C.c_int, _dp]
        L.orc_flux.restype = C.c_double
        L.orc_image2vis.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_image2vis.restype = C.c_double
        L.orc_vis2data.argtypes = [C.c_int, C.c_int, _dp, _ip, _ip, _dp, _dp]
        L.orc_data2chi2.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_data2chi2.restype = C.c_double
        L.orc_entropy.argtypes = [C.c_int, _dp, _dp]
        L.orc_entropy.restype = C.c_double
        L.orc_entropy_grad.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_grad_direct.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp]
        L.orc_grad_coeffs.argtypes = [C.c_int] * 3 + [_dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp]
        L.orc_grad_coeffs.restype = C.c_double
        L.orc_grad_separable.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_cg.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp, C.c_double, C.c_int,
                                             C.c_int, C.c_double, _dp, C.c_void_p]
        L.orc_cg.restype = C.c_int
        L.orc_pcg.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp, C.c_double, C.c_int,
                                              C.c_void_p, C.c_int, C.c_double, _dp, C.c_void_p]
        L.orc_pcg.restype = C.c_int
        L.orc_dft_raw.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_parabola_step.argtypes = [C.c_double] * 6
        L.orc_px_update.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, _ip, _dp, _dp]
        L.orc_px_update.restype = C.c_double
        L.orc_px_propose.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, _dp, _ip, _ip, _dp, _dp, _dp, _dp,
                                                     C.c_double, C.c_int, _ip, _ip, _dp, _dp, _dp, _dp]
        L.orc_parabola_step.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_linesearch.argtypes = [PHI, PHI, C.c_void_p, C.c_double, C.c_double, C.c_int, _dp]
        L.orc_linesearch.restype = C.c_double
        L.orc_ls_zoom.argtypes = [PHI, PHI, C.c_void_p] + [C.c_double] * 5 + [C.c_int, _dp]
        L.orc_ls_zoom.restype = C.c_double
        L.orc_set_plant.argtypes = [C.c_int]
        L.orc_set_ls_log.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_ls_cap.argtypes = [C.c_int]

    # -- problem helpers ----------------------------------------------------
    @staticmethod
    def unpack(prob):
        """problem (gpair_amd.synth.Problem) -> FP64 arrays used by every call."""
        return dict(W=prob.W, nuv=prob.nuv, npow=prob.npow, nbis=prob.nbis, au=_c64(prob.au()),
                    av=_c64(prob.av()), data=_c64(prob.data), err=_c64(prob.err),
                    phasor=_c64(np.stack([prob.phasor.real, prob.phasor.imag], 1).ravel()),
                    uvpnt=_i32(prob.uvpnt.ravel()), sign=_i32(prob.sign.ravel()), model=_c64(prob.model))

    def threads(self) -> int:
        return self.L.orc_num_threads()

    def tables(self, W, au, av):
        nuv = au.size
        X = np.empty(2 * nuv * W)
        Y = np.empty(2 * nuv * W)
        self.L.orc_tables(W, nuv, _c64(au), _c64(av), X, Y)
        return X.view(np.complex128).reshape(nuv, W), Y.view(np.complex128).reshape(nuv, W)

    def flux(self, img):
        img = _c64(img)
        return self.L.orc_flux(img.size, img)

    def image2vis(self, W, au, av, img):
        vis = np.empty(2 * au.size)
        F = self.L.orc_image2vis(W, au.size, _c64(au), _c64(av), _c64(img), vis)
        return vis.view(np.complex128), F

    def vis2data(self, npow, nbis, vis, uvpnt, sign, phasor):
        mock = np.empty(npow + 2 * nbis)
        self.L.orc_vis2data(npow, nbis, _c64(np.asarray(vis, np.complex128).view(np.float64)), _i32(uvpnt),
                            _i32(sign), _c64(phasor), mock)
        return mock

    def chi2(self, data, err, mock):
        return self.L.orc_data2chi2(mock.size, _c64(data), _c64(err), _c64(mock))

    def entropy(self, img, model):
        return self.L.orc_entropy(img.size, _c64(img), _c64(model))

    def entropy_grad(self, img, model, old=None):
        g = np.zeros(img.size) if old is None else _c64(old).copy()
        self.L.orc_entropy_grad(img.size, _c64(img), _c64(model), g)
        return g

    def grad_direct(self, P, img, vis, mock):
        g = np.empty(P["W"] ** 2)
        self.L.orc_grad_direct(P["W"], P["nuv"], P["npow"], P["nbis"], P["au"], P["av"], _c64(img),
                               _c64(np.asarray(vis, np.complex128).view(np.float64)), _c64(mock),
                               P["data"], P["err"], P["uvpnt"], P["sign"], P["phasor"], g)
        return g

    def grad_coeffs(self, P, vis, mock):
        coef = np.empty(2 * P["nuv"])
        c0 = self.L.orc_grad_coeffs(P["nuv"], P["npow"], P["nbis"],
                                    _c64(np.asarray(vis, np.complex128).view(np.float64)), _c64(mock),
                                    P["data"], P["err"], P["uvpnt"], P["sign"], P["phasor"], coef)
        return coef.view(np.complex128), c0

    def grad_separable(self, P, coef, c0, F):
        g = np.empty(P["W"] ** 2)
        self.L.orc_grad_separable(P["W"], P["nuv"], P["au"], P["av"],
                                  _c64(np.asarray(coef, np.complex128).view(np.float64)), c0, F, g)
        return g

    # -- composite ------------------------------------------------------------
    def objective(self, P, img, alpha):
        """J, chi2, S, vis, mock, F for an image (FP64)."""
        vis, F = self.image2vis(P["W"], P["au"], P["av"], img)
        mock = self.vis2data(P["npow"], P["nbis"], vis, P["uvpnt"], P["sign"], P["phasor"])
        chi2 = self.chi2(P["data"], P["err"], mock)
        S = self.entropy(img, P["model"])
        return chi2 - alpha * S, chi2, S, vis, mock, F

    def gradient(self, P, img, alpha, vis, mock):
        """full criterion gradient (separable form) and its parts."""
        coef, c0 = self.grad_coeffs(P, vis, mock)
        gd = self.grad_separable(P, coef, c0, self.flux(img))
        ge = self.entropy_grad(img, P["model"])
        return gd - alpha * ge, gd, ge

    def cg(self, P, img, alpha, method=2, max_iter=200, chi2r_target=1.0, ls_log=False):
        """reference-rule loop; with ls_log also the line-search record of every
        iteration: rows of (exit, s, J(s), w2, nJ, nG, counter, width, w1, J0)"""
        img = _c64(img).copy()
        trace = np.zeros((max_iter, 8))
        log = np.full((max_iter, 10), np.nan) if ls_log else None
        if ls_log:
            self.L.orc_set_ls_log(log.ctypes.data, max_iter)
        try:
            rows = self.L.orc_cg(P["W"], P["nuv"], P["npow"], P["nbis"], P["au"], P["av"], P["data"], P["err"],
                                 P["uvpnt"], P["sign"], P["phasor"], P["model"], float(alpha), int(method),
                                 int(max_iter), float(chi2r_target), img, trace.ctypes.data)
        finally:
            if ls_log:
                self.L.orc_set_ls_log(None, 0)
        if ls_log:
            return img, trace[:rows], log[:rows]
        return img, trace[:rows]

    # -- the reference line search on a 1-D objective ------------------------
    def _ls(self, fn, phi, dphi, *args, cap=64):
        calls = []

        def _phi(_u, s):
            calls.append(("J", s))
            return float(phi(s))

        def _dphi(_u, s):
            calls.append(("g", s))
            return float(dphi(s))
        info = np.zeros(8)
        f1, f2 = PHI(_phi), PHI(_dphi)   # kept alive for the call
        s = fn(f1, f2, None, *[float(a) for a in args], int(cap), info)
        return s, dict(zip(LS_INFO, info)), calls

    def linesearch(self, phi, dphi, J0, w1, cap=64):
        """bracket + zoom (gpair.c:791-906, gpu.c:1947-2065) on phi(s); returns
        the step, the info dict and the list of evaluations ("J" | "g", s)"""
        return self._ls(self.L.orc_linesearch, phi, dphi, J0, w1, cap=cap)

    def zoom(self, phi, dphi, lo, hi, J_lo, w1, J0, cap=64):
        """the zoom alone (gpu_linesearch_zoom's arguments)"""
        return self._ls(self.L.orc_ls_zoom, phi, dphi, lo, hi, J_lo, w1, J0, cap=cap)

    def plant(self, mask):
        """planted mistakes in the zoom (tests only): 1 strong curvature test,
        2 J_lo updated when lo moves, 4 bisection only; 0 restores the rules"""
        self.L.orc_set_plant(int(mask))

    def pcg(self, P, img, alpha, precond=1, pfactor=None, max_iter=200, chi2r_target=1.0):
        """accelerated loop (preconditioned PR-CG, batched line search)."""
        img = _c64(img).copy()
        trace = np.zeros((max_iter + 1, 8))
        pf = None if pfactor is None else _c64(pfactor)
        rows = self.L.orc_pcg(P["W"], P["nuv"], P["npow"], P["nbis"], P["au"], P["av"], P["data"], P["err"],
                              P["uvpnt"], P["sign"], P["phasor"], P["model"], float(alpha), int(precond),
                              None if pf is None else pf.ctypes.data, int(max_iter), float(chi2r_target), img,
                              trace.ctypes.data)
        return img, trace[:rows]

    def px_update(self, P, vis, F, pix, d):
        """visibilities and flux after pixel changes d[j] at pix[j] (x + W*y)"""
        pix = _i32(np.atleast_1d(pix))
        d = _c64(np.atleast_1d(d))
        out = np.empty(2 * P["nuv"])
        Fn = self.L.orc_px_update(P["W"], P["nuv"], P["au"], P["av"],
                                  _c64(np.asarray(vis, np.complex128).view(np.float64)), float(F), pix.size, pix, d,
                                  out)
        return out.view(np.complex128), Fn

    def px_propose(self, P, img, vis, F, pix_a, pix_b, flux_a, flux_b):
        """dchi2, dS of each move (pixels pix_a / pix_b set to flux_a / flux_b; pix_b = -1: none)"""
        pix_a, pix_b = _i32(pix_a), _i32(pix_b)
        k = pix_a.size
        dchi2, dS = np.empty(k), np.empty(k)
        self.L.orc_px_propose(P["W"], P["nuv"], P["npow"], P["nbis"], P["au"], P["av"], P["data"], P["err"],
                              P["uvpnt"], P["sign"], P["phasor"], P["model"], _c64(img),
                              _c64(np.asarray(vis, np.complex128).view(np.float64)), float(F), k, pix_a, pix_b,
                              _c64(flux_a), _c64(flux_b), dchi2, dS)
        return dchi2, dS

    def dft_raw(self, W, au, av, img):
        out = np.empty(2 * au.size)
        self.L.orc_dft_raw(W, au.size, _c64(au), _c64(av), _c64(img), out)
        return out.view(np.complex128)


# ---------------------------------------------------------------------------
# the reference's own CPU model (gpair1/trunk/cpu.c) through ctypes
# ---------------------------------------------------------------------------
class _UvPnt(C.Structure):          # gpair1/trunk/getoifits.h:37-41
    _fields_ = [("sign", C.c_short), ("uvpnt", C.c_long)]


class _BsRef(C.Structure):          # gpair1/trunk/getoifits.h:43-51
    _fields_ = [("ab", _UvPnt), ("bc", _UvPnt), ("ca", _UvPnt)]


class _Uv(C.Structure):             # gpair1/trunk/getoifits.h:53-59
    _fields_ = [("u", C.c_float), ("v", C.c_float), ("bandwidth", C.c_float), ("wavelength", C.c_float)]


class _OiData(C.Structure):         # gpair1/trunk/getoifits.h:61-75
    _fields_ = [("pow", C.c_void_p), ("powerr", C.c_void_p), ("bisamp", C.c_void_p), ("bisphs", C.c_void_p),
                ("bisamperr", C.c_void_p), ("bisphserr", C.c_void_p), ("time", C.c_void_p),
                ("uv", C.POINTER(_Uv)), ("bsref", C.POINTER(_BsRef)), ("npow", C.c_int), ("nbis", C.c_int),
                ("nuv", C.c_int)]


class _Chi2Info(C.Structure):       # gpair1/trunk/data_types.h:10-32
    _fields_ = [("npow", C.c_int), ("nbis", C.c_int), ("nuv", C.c_int), ("data", C.c_void_p),
                ("data_err", C.c_void_p), ("data_phasor", C.c_void_p), ("oifits_info", C.POINTER(_OiData)),
                ("image_width", C.c_int), ("dft_x", C.c_void_p), ("dft_y", C.c_void_p), ("visi", C.c_void_p),
                ("mock", C.c_void_p)]


class RefCPU:
    """Float-precision reference model; arrays are kept alive on the instance."""

    def __init__(self, prob, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = C.CDLL(path)
        self.L = L
        self.prob = prob
        W, nuv, npow, nbis = prob.W, prob.nuv, prob.npow, prob.nbis
        self.W = W
        # DFT tables [nuv][W] float complex, gpair1/trunk/gpair.c:231-245
        X, Y = Oracle().tables(W, prob.au(), prob.av())
        self.X = np.ascontiguousarray(X.astype(np.complex64))
        self.Y = np.ascontiguousarray(Y.astype(np.complex64))
        self.data = np.ascontiguousarray(prob.data, np.float32)
        self.err = np.ascontiguousarray(prob.err, np.float32)
        self.phasor = np.ascontiguousarray(prob.phasor, np.complex64)
        self.visi = np.zeros(nuv, np.complex64)
        self.mock = np.zeros(npow + 2 * nbis, np.float32)
        self.bsref = (_BsRef * max(nbis, 1))()
        for t in range(nbis):
            for j, name in enumerate(("ab", "bc", "ca")):
                e = getattr(self.bsref[t], name)
                e.sign = int(prob.sign[t, j])
                e.uvpnt = int(prob.uvpnt[t, j])
        self.uv = (_Uv * nuv)()
        for p in range(nuv):
            self.uv[p].u = float(prob.u[p])
            self.uv[p].v = float(prob.v[p])
        self.oi = _OiData()
        self.oi.uv = self.uv
        self.oi.bsref = self.bsref
        self.oi.npow, self.oi.nbis, self.oi.nuv = npow, nbis, nuv
        self.info = _Chi2Info()
        self.info.npow, self.info.nbis, self.info.nuv = npow, nbis, nuv
        self.info.data = self.data.ctypes.data
        self.info.data_err = self.err.ctypes.data
        self.info.data_phasor = self.phasor.ctypes.data
        self.info.oifits_info = C.pointer(self.oi)
        self.info.image_width = W
        self.info.dft_x = self.X.ctypes.data
        self.info.dft_y = self.Y.ctypes.data
        self.info.visi = self.visi.ctypes.data
        self.info.mock = self.mock.ctypes.data
        L.image2chi2.argtypes = [C.POINTER(_Chi2Info), C.c_void_p]
        L.image2chi2.restype = C.c_float
        L.compute_data_gradient.argtypes = [C.POINTER(_Chi2Info), C.c_void_p, C.c_void_p]
        L.GullSkilling_entropy.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.GullSkilling_entropy.restype = C.c_float
        L.GullSkilling_entropy_gradient.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.set_model.argtypes = [C.c_int, C.c_float, C.c_int, C.c_float, C.c_float, C.c_void_p]

    def image2chi2(self, img):
        """chi2 (float) and copies of the visibilities and mock data, cpu.c:435-447."""
        img = np.ascontiguousarray(img, np.float32)
        chi2 = self.L.image2chi2(C.byref(self.info), img.ctypes.data)
        return float(chi2), self.visi.copy(), self.mock.copy()

    def data_gradient(self, img):
        """cpu.c:279-369; needs visi/mock of img (call image2chi2 first)."""
        img = np.ascontiguousarray(img, np.float32)
        g = np.zeros(self.W * self.W, np.float32)
        self.L.compute_data_gradient(C.byref(self.info), img.ctypes.data, g.ctypes.data)
        return g

    def entropy(self, img, model):
        img = np.ascontiguousarray(img, np.float32)
        model = np.ascontiguousarray(model, np.float32)
        return float(self.L.GullSkilling_entropy(self.W, img.ctypes.data, model.ctypes.data))

    def entropy_grad(self, img, model, old=None):
        img = np.ascontiguousarray(img, np.float32)
        model = np.ascontiguousarray(model, np.float32)
        g = np.zeros(self.W * self.W, np.float32) if old is None else np.array(old, np.float32)
        self.L.GullSkilling_entropy_gradient(self.W, img.ctypes.data, model.ctypes.data, g.ctypes.data)
        return g

    def set_model(self, pix, modeltype, width, flux):
        out = np.zeros(self.W * self.W, np.float32)
        self.L.set_model(self.W, pix, modeltype, width, flux, out.ctypes.data)
        return out


def ref_available() -> bool:
    return os.path.exists(REF_SO)


REF_OIFITS = os.path.join(HERE, "_ref", "ref_oifits")
REF_LOOP = os.path.join(HERE, "_ref", "gpair_cpu")


def ref_oifits_read(path, target_id=-1, minband=-1.0, maxband=-1.0, workdir=None):
    """the reference's own loader (get_oi_fits_selection + get_oi_fits_data,
    gpair1/trunk/getoifits.c, compiled into oracle/_ref/ref_oifits): a dict of
    arrays in the keys of oifits_oracle.read.  The band defaults as the product's
    (1 .. 50000 nm); a target id goes to the reference's prompt, which it reads
    only for files with several targets."""
    import tempfile
    if minband < 0 or maxband <= minband:
        minband, maxband = 1.0, 50000.0
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        out = os.path.join(d, "oi.bin")
        stdin = f"{target_id}\n" if target_id >= 0 else ""
        r = subprocess.run([REF_OIFITS, os.path.abspath(path), out, repr(float(minband)), repr(float(maxband))], input=stdin,
                           capture_output=True, text=True, timeout=60, cwd=d)
        if r.returncode != 0:
            raise RuntimeError(f"ref_oifits failed ({r.returncode}): {r.stderr[-500:]}")
        raw = open(out, "rb").read()
    npow, nbis, nuv = np.frombuffer(raw[:12], np.int32)
    o = 12
    res = {}

    def take(name, dtype, count, shape=None):
        nonlocal o
        a = np.frombuffer(raw, dtype, count, o).copy()
        o += a.nbytes
        res[name] = a.reshape(shape) if shape else a
    for k in ("pow", "powerr"):
        take(k, np.float32, npow)
    for k in ("bisamp", "bisamperr", "bisphs", "bisphserr"):
        take(k, np.float32, nbis)
    for k in ("u", "v", "wavelength", "bandwidth"):
        take(k, np.float32, nuv)
    take("uvpnt", np.int64, 3 * nbis, (-1, 3))
    take("sign", np.int16, 3 * nbis, (-1, 3))
    assert o == len(raw)
    return res
