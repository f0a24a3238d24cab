"""OIFITS and FITS image I/O over the C ABI (include/gpair_io.h).

``read`` replaces the reference's read_oifits (gpair1/trunk/gpair.c:944-951:
get_oi_fits_selection + get_oi_fits_data) without its prompts; ``problem``
chains it with init_data and set_model (gpair1/trunk/gpair.c:997-1066,
gpair1/trunk/cpu.c:645-749) into the packed arrays the GPU context takes;
``write`` / ``write_problem`` produce OIFITS files; ``read_image`` /
``write_image`` replace read_fits_image / writefits (gpair1/trunk/gpair.c:954-995,
1164-1194).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib

PI_REF = 3.14159265358979323


class OifitsError(RuntimeError):
    pass


def _L():
    return _lib.load()


def _check(rc, what):
    if rc != 0:
        msg = _L().gpair_io_last_error()
        raise OifitsError(f"{what}: {msg.decode() if msg else rc}")


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclasses.dataclass
class OIData:
    """gpair_oidata as numpy arrays (see include/gpair_io.h)"""
    pow: np.ndarray
    powerr: np.ndarray
    bisamp: np.ndarray
    bisamperr: np.ndarray
    bisphs: np.ndarray
    bisphserr: np.ndarray
    u: np.ndarray
    v: np.ndarray
    wavelength: np.ndarray
    bandwidth: np.ndarray
    uvpnt: np.ndarray      # int64 [nbis, 3]
    sign: np.ndarray       # int16 [nbis, 3]
    target_id: int
    minband: float
    maxband: float
    uv_compares: int

    @property
    def npow(self):
        return int(self.pow.size)

    @property
    def nbis(self):
        return int(self.bisamp.size)

    @property
    def nuv(self):
        return int(self.u.size)


def targets(path):
    """[(target_id, name)] of an OIFITS file"""
    n = C.c_int()
    _check(_L().gpair_oifits_targets(path.encode(), None, None, 0, C.byref(n)), "gpair_oifits_targets")
    ids = np.zeros(max(n.value, 1), np.int32)
    names = C.create_string_buffer(17 * max(n.value, 1))
    _check(_L().gpair_oifits_targets(path.encode(), _p(ids), names, n.value, C.byref(n)), "gpair_oifits_targets")
    return [(int(ids[i]), names.raw[17 * i:17 * (i + 1)].split(b"\0")[0].decode()) for i in range(n.value)]


def read(path, target_id=-1, minband=-1.0, maxband=-1.0) -> OIData:
    d = _lib.OiData()
    _check(_L().gpair_oifits_read(path.encode(), int(target_id), float(minband), float(maxband), C.byref(d)),
           "gpair_oifits_read")
    try:
        def arr(ptr, n, dt, shape=None):
            if n == 0:
                return np.zeros(shape or (0,), dt)
            a = np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)
            return a.reshape(shape) if shape else a
        out = OIData(pow=arr(d.pow, d.npow, np.float32), powerr=arr(d.powerr, d.npow, np.float32),
                     bisamp=arr(d.bisamp, d.nbis, np.float32), bisamperr=arr(d.bisamperr, d.nbis, np.float32),
                     bisphs=arr(d.bisphs, d.nbis, np.float32), bisphserr=arr(d.bisphserr, d.nbis, np.float32),
                     u=arr(d.u, d.nuv, np.float32), v=arr(d.v, d.nuv, np.float32),
                     wavelength=arr(d.wavelength, d.nuv, np.float32), bandwidth=arr(d.bandwidth, d.nuv, np.float32),
                     uvpnt=arr(d.uvpnt, 3 * d.nbis, np.int64, (d.nbis, 3)),
                     sign=arr(d.sign, 3 * d.nbis, np.int16, (d.nbis, 3)), target_id=d.target_id,
                     minband=d.minband, maxband=d.maxband, uv_compares=d.uv_compares)
    finally:
        _L().gpair_oidata_free(C.byref(d))
    return out


def write(path, eff_wave, eff_band, v2_u, v2_v, vis2data, vis2err, t3_u1, t3_v1, t3_u2, t3_v2, t3amp, t3amperr,
          t3phi, t3phierr, v2_flag=None, t3_flag=None, target="SYNTH"):
    """OIFITS v1 file: coordinates in metres, data arrays [record, channel]"""
    ew = np.ascontiguousarray(eff_wave, np.float32)
    eb = np.ascontiguousarray(eff_band, np.float32)
    nw = ew.size

    def d(a, shape):
        return np.ascontiguousarray(np.asarray(a, np.float64).reshape(shape))
    nv = int(np.asarray(v2_u).size)
    nt = int(np.asarray(t3_u1).size)
    vf = None if v2_flag is None else np.ascontiguousarray(np.asarray(v2_flag, np.uint8).reshape(nv, nw))
    tf = None if t3_flag is None else np.ascontiguousarray(np.asarray(t3_flag, np.uint8).reshape(nt, nw))
    args = [d(v2_u, nv), d(v2_v, nv), d(vis2data, (nv, nw)), d(vis2err, (nv, nw)), vf,
            d(t3_u1, nt), d(t3_v1, nt), d(t3_u2, nt), d(t3_v2, nt), d(t3amp, (nt, nw)), d(t3amperr, (nt, nw)),
            d(t3phi, (nt, nw)), d(t3phierr, (nt, nw)), tf]
    p = [_p(a) for a in args]
    _check(_L().gpair_oifits_write(path.encode(), target.encode(), nw, _p(ew), _p(eb), nv, *p[:5], nt, *p[5:]),
           "gpair_oifits_write")


def write_problem(path, prob, wave0=1.0e-6):
    """A gpair_amd.synth problem as OIFITS: channel c of the layout observes at
    wave0 * (1 + 0.6 c / (nch - 1)) (the layout's relative wavelengths), every
    snapshot / baseline is one OI_VIS2 record and every snapshot / triangle one
    OI_T3 record over all channels, with coordinates in metres.  Noise sigmas as
    the problem's (V2 error from the inverse errors, T3 amplitude error, 1 degree
    phase error); phases in degrees."""
    m = prob.meta
    nsnap, nch, nbase, ntri = m["nsnap"], m["nch"], m["nbase"], m["ntri"]
    if prob.npow != nsnap * nch * nbase or prob.nbis != nsnap * nch * ntri:
        raise OifitsError("write_problem needs a problem with every baseline and triangle (no T3 cut)")
    lam = wave0 * np.linspace(1.0, 1.6, nch)
    u = prob.u.astype(np.float64).reshape(nsnap, nch, nbase)
    v = prob.v.astype(np.float64).reshape(nsnap, nch, nbase)
    # stored points are flipped into +u; channel 0 fixes the record's baseline in metres
    bu, bv = u[:, 0, :] * lam[0], v[:, 0, :] * lam[0]
    v2 = prob.data[:prob.npow].astype(np.float64).reshape(nsnap, nch, nbase).transpose(0, 2, 1)
    v2e = (1.0 / prob.err[:prob.npow].astype(np.float64)).reshape(nsnap, nch, nbase).transpose(0, 2, 1)
    # T3 legs ab, bc of each triangle as metres, from the layout's uv points and signs
    up = prob.uvpnt.reshape(nsnap, nch, ntri, 3)[:, 0]
    sg = prob.sign.reshape(nsnap, nch, ntri, 3)[:, 0].astype(np.float64)
    uf = prob.u.astype(np.float64) * lam[0]
    vf = prob.v.astype(np.float64) * lam[0]
    u1, v1 = sg[..., 0] * uf[up[..., 0]], sg[..., 0] * vf[up[..., 0]]
    u2, v2_ = sg[..., 1] * uf[up[..., 1]], sg[..., 1] * vf[up[..., 1]]
    amp = prob.data[prob.npow::2].astype(np.float64).reshape(nsnap, nch, ntri).transpose(0, 2, 1)
    ampe = (1.0 / prob.err[prob.npow::2].astype(np.float64)).reshape(nsnap, nch, ntri).transpose(0, 2, 1)
    phi = -np.degrees(np.angle(prob.phasor.astype(np.complex128))).reshape(nsnap, nch, ntri).transpose(0, 2, 1)
    phie = np.ones_like(phi)
    write(path, lam, 0.01 * lam, bu.ravel(), bv.ravel(), v2.reshape(-1, nch), v2e.reshape(-1, nch),
          u1.ravel(), v1.ravel(), u2.ravel(), v2_.ravel(), amp.reshape(-1, nch), ampe.reshape(-1, nch),
          phi.reshape(-1, nch), phie.reshape(-1, nch))


def pack(oi: OIData, do_extrapolation=1):
    """init_data (gpair1/trunk/gpair.c:997-1066) through gpair_init_data:
    data, inverse errors, phasors (complex64) and ndof"""
    L = _L()
    npow, nbis = oi.npow, oi.nbis
    data = np.zeros(npow + 2 * nbis, np.float32)
    err = np.zeros(npow + 2 * nbis, np.float32)
    ph = np.zeros(2 * max(nbis, 1), np.float32)
    bamp = oi.bisamp.copy()
    bampe = oi.bisamperr.copy()
    ndof = C.c_int()
    uvp = np.ascontiguousarray(oi.uvpnt.reshape(-1), np.int64)
    _check(L.gpair_init_data(npow, nbis, _p(oi.pow), _p(oi.powerr), _p(bamp), _p(bampe), _p(oi.bisphs),
                             _p(oi.bisphserr), _p(uvp), int(do_extrapolation), _p(data), _p(err), _p(ph),
                             C.byref(ndof)), "gpair_init_data")
    return data, err, ph[:2 * nbis].view(np.complex64), ndof.value


def problem(path, W, pixellation, model_type=0, model_width=0.0, model_flux=None, target_id=-1, minband=-1.0,
            maxband=-1.0, do_extrapolation=1, centered=True, init_image=None, model_image=None):
    """A gpair_amd.synth.Problem read from an OIFITS file (the reference's set-up,
    gpair1/trunk/gpair.c:150-229): data packed by init_data, prior from set_model
    or a FITS image, start = max(prior, 1e-8) or a FITS image."""
    from . import synth
    oi = read(path, target_id, minband, maxband)
    data, err, phasor, ndof = pack(oi, do_extrapolation)
    n = W * W
    if model_image is not None:
        model = read_image(model_image).astype(np.float32).ravel()
        if model.size != n:
            raise OifitsError(f"prior image {model_image}: {model.size} pixels, expected {n}")
    else:
        flux = float(n) if model_flux is None else float(model_flux)
        model = np.zeros(n, np.float32)
        _check(_L().gpair_set_model(_p(model), W, float(pixellation), int(model_type), float(model_width), flux,
                                    int(bool(centered))), "gpair_set_model")
    if init_image is not None:
        start = read_image(init_image).astype(np.float32).ravel()
        if start.size != n:
            raise OifitsError(f"start image {init_image}: {start.size} pixels, expected {n}")
    else:
        start = np.maximum(model, np.float32(1e-8))
    meta = dict(cfg="oifits", path=path, target_id=oi.target_id, band=(oi.minband, oi.maxband), ndof_init=ndof,
                alpha=None, prior=("file" if model_image else model_type, model_width))
    return synth.Problem(W=W, pixellation=pixellation, u=oi.u.copy(), v=oi.v.copy(), npow=oi.npow, nbis=oi.nbis,
                         data=data, err=err, phasor=phasor, uvpnt=oi.uvpnt.copy(), sign=oi.sign.copy(), model=model,
                         start=start, truth_vis=None, meta=meta)


def read_image(path):
    """primary image as float32 [height, width]"""
    w, h = C.c_int(), C.c_int()
    _check(_L().gpair_fits_read_image(path.encode(), None, 0, C.byref(w), C.byref(h)), "gpair_fits_read_image")
    out = np.empty((h.value, w.value), np.float32)
    _check(_L().gpair_fits_read_image(path.encode(), _p(out), out.size, C.byref(w), C.byref(h)),
           "gpair_fits_read_image")
    return out


def write_image(path, image, pixsize=0.0, bitpix=-64, datafile=None, target=None):
    img = np.ascontiguousarray(image, np.float32)
    if img.ndim == 1:
        W = int(round(img.size ** 0.5))
        img = img.reshape(W, W)
    h, w = img.shape
    _check(_L().gpair_fits_write_image(path.encode(), _p(img), w, h, int(bitpix), float(pixsize),
                                       None if datafile is None else datafile.encode(),
                                       None if target is None else target.encode()), "gpair_fits_write_image")
