"""Python host side over the C ABI.

``GpairContext`` keeps the reference's call surface (gpair1/trunk/gpu.h): the
methods carry the reference names (``gpu_get_chi2_curr``, ``gpu_linesearch_zoom``,
...) with the same argument meaning; errors raise ``GpairError`` instead of the
reference's exit(0) (gpair1/trunk/gpu.c:186-202).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

PREC_FP32 = 0
PREC_BF16X3 = 1
PREC_FP16S = 2


class GpairError(RuntimeError):
    pass


# gpair_pixel_move (include/gpair_hip.h): pixels x + W*y, pix_b = -1 for one pixel
PIXEL_MOVE = np.dtype([("pix_a", np.int32), ("pix_b", np.int32), ("flux_a", np.float32), ("flux_b", np.float32)])


def pixel_moves(moves):
    """moves as a contiguous PIXEL_MOVE array; accepts such an array or a
    sequence of (pix_a, pix_b, flux_a, flux_b) tuples"""
    if isinstance(moves, np.ndarray) and moves.dtype == PIXEL_MOVE:
        return np.ascontiguousarray(moves)
    return np.array([tuple(m) for m in moves], dtype=PIXEL_MOVE)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class GpairContext:
    def __init__(self, device: int = 0):
        self.L = _lib.load()
        h = C.c_void_p()
        rc = self.L.gpu_init(C.byref(h), device)
        if rc != 0:
            raise GpairError(f"gpu_init: {self.L.gpair_strerror(rc).decode()}")
        self.h = h
        self.W = self.nuv = self.npow = self.nbis = 0

    # -- plumbing -----------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = self.L.gpair_last_error(self.h)
            raise GpairError(f"{what}: {self.L.gpair_strerror(rc).decode()}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.L.gpu_cleanup(self.h)
            self.h = None

    gpu_cleanup = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ndof(self):
        return self.npow + 2 * self.nbis

    # -- setup (gpair1/trunk/gpu.c:1175-1451) ----------------------------------
    def gpu_copy_data(self, data, data_err, nuv, phasor, npow, uvpnt4, sign4, default_model, W):
        data = np.ascontiguousarray(data, np.float32)
        err = np.ascontiguousarray(data_err, np.float32)
        ph = np.ascontiguousarray(np.asarray(phasor, np.complex64)).view(np.float32)
        uv4 = np.ascontiguousarray(uvpnt4, np.int64)
        s4 = np.ascontiguousarray(sign4, np.int16)
        model = np.ascontiguousarray(default_model, np.float32)
        nbis = uv4.shape[0]
        self._check(self.L.gpu_copy_data(self.h, _ptr(data), _ptr(err), data.size, int(nuv), _ptr(ph), nbis,
                                         int(npow), _ptr(uv4), _ptr(s4), nbis, _ptr(model), model.size, int(W)),
                    "gpu_copy_data")
        self.W, self.nuv, self.npow, self.nbis = int(W), int(nuv), int(npow), int(nbis)

    def gpu_copy_image(self, image):
        img = np.ascontiguousarray(image, np.float32)
        self._check(self.L.gpu_copy_image(self.h, _ptr(img), self.W, self.W), "gpu_copy_image")

    def gpair_set_uv(self, u, v, pixellation):
        u = np.ascontiguousarray(u, np.float32)
        v = np.ascontiguousarray(v, np.float32)
        self._check(self.L.gpair_set_uv(self.h, _ptr(u), _ptr(v), u.size, float(pixellation)), "gpair_set_uv")

    def gpu_copy_dft(self, dft_x, dft_y):
        """tables in the reference GPU layout, complex [W][nuv]."""
        x = np.ascontiguousarray(np.asarray(dft_x, np.complex64))
        y = np.ascontiguousarray(np.asarray(dft_y, np.complex64))
        self._check(self.L.gpu_copy_dft(self.h, _ptr(x), _ptr(y), x.size), "gpu_copy_dft")

    def gpu_build_kernels(self):
        self._check(self.L.gpu_build_kernels(self.h, self.ndof, self.W, self.W * self.W), "gpu_build_kernels")

    def set_precision(self, prec: int):
        self._check(self.L.gpair_set_precision(self.h, int(prec)), "gpair_set_precision")

    @classmethod
    def from_problem(cls, prob, device: int = 0, image=None):
        """full set-up from a gpair_amd.synth.Problem (tables built on the device)."""
        ctx = cls(device)
        ctx.gpu_copy_data(prob.data, prob.err, prob.nuv, prob.phasor, prob.npow, prob.uvpnt4(), prob.sign4(),
                          prob.model, prob.W)
        ctx.gpair_set_uv(prob.u, prob.v, prob.pixellation)
        ctx.gpu_build_kernels()
        ctx.gpu_copy_image(prob.start if image is None else image)
        return ctx

    # -- buffers --------------------------------------------------------------
    def buf(self, name):
        f = {"ci": self.L.gpu_getp_ci, "ti": self.L.gpu_getp_ti, "fgn": self.L.gpu_getp_fgn,
             "fg": self.L.gpu_getp_fg, "dd": self.L.gpu_getp_dd, "tg": self.L.gpu_getp_tg,
             "eg": self.L.gpu_getp_eg, "dg": self.L.gpair_getp_dg}[name]
        return C.c_void_p(f(self.h))

    def get(self, name):
        out = np.empty(self.W * self.W, np.float32)
        self._check(self.L.gpu_get_image(self.h, out.size, _ptr(out), self.buf(name)), "gpu_get_image")
        return out

    def put(self, name, arr):
        a = np.ascontiguousarray(arr, np.float32)
        self._check(self.L.gpair_put_image(self.h, self.buf(name), _ptr(a), a.size), "gpair_put_image")

    def vis(self):
        out = np.empty(self.nuv, np.complex64)
        self._check(self.L.gpair_get_vis(self.h, _ptr(out), self.nuv), "gpair_get_vis")
        return out

    def mock(self):
        out = np.empty(self.ndof, np.float32)
        self._check(self.L.gpair_get_mock(self.h, _ptr(out), self.ndof), "gpair_get_mock")
        return out

    # -- single-pixel moves (gpair1/trunk/gpu.c:2067-2182, gpair1/trunk/cpu.c:477-497)
    def pixel_begin(self):
        """(re)build the pixel state from the current image (one forward DFT)"""
        self._check(self.L.gpair_pixel_begin(self.h), "gpair_pixel_begin")

    def gpu_update_vis_fluxchange(self, x, y, new_pixel_flux):
        """trial visibilities of the pixel state with pixel (x, y) set to new_pixel_flux"""
        self._check(self.L.gpu_update_vis_fluxchange(self.h, int(x), int(y), float(new_pixel_flux), self.W,
                                                     self.nuv, self.nuv), "gpu_update_vis_fluxchange")

    def update_vis_positionchange(self, x_old, y_old, x_new, y_new, flux):
        """trial visibilities with `flux` moved from (x_old, y_old) to (x_new, y_new)"""
        self._check(self.L.gpair_update_vis_positionchange(self.h, int(x_old), int(y_old), int(x_new), int(y_new),
                                                           float(flux), self.W, self.nuv),
                    "gpair_update_vis_positionchange")

    def gpu_new_chi2(self):
        """chi2 of the trial visibilities"""
        v = C.c_double()
        self._check(self.L.gpu_new_chi2(self.h, self.nuv, self.npow, self.nbis, self.ndof, C.byref(v)),
                    "gpu_new_chi2")
        return v.value

    def trial_vis(self):
        out = np.empty(self.nuv, np.complex64)
        self._check(self.L.gpair_get_trial_vis(self.h, _ptr(out), self.nuv), "gpair_get_trial_vis")
        return out

    def pixel_propose(self, moves):
        """dchi2, dentropy (FP64 arrays) of each move against the pixel state;
        moves: a PIXEL_MOVE structured array (or anything pixel_moves accepts)"""
        m = pixel_moves(moves)
        dchi2 = np.empty(m.size, np.float64)
        dS = np.empty(m.size, np.float64)
        self._check(self.L.gpair_pixel_propose(self.h, _ptr(m), m.size, _ptr(dchi2), _ptr(dS)),
                    "gpair_pixel_propose")
        return dchi2, dS

    def pixel_apply(self, moves):
        m = pixel_moves(moves)
        self._check(self.L.gpair_pixel_apply(self.h, _ptr(m), m.size), "gpair_pixel_apply")

    def pixel_state(self):
        """(chi2, entropy, flux) of the pixel state"""
        c2, S, F = C.c_double(), C.c_double(), C.c_double()
        self._check(self.L.gpair_pixel_state(self.h, C.byref(c2), C.byref(S), C.byref(F)), "gpair_pixel_state")
        return c2.value, S.value, F.value

    # -- objective / gradients (gpair1/trunk/gpu.c:596-857, 1598-1802) ---------
    def gpu_get_chi2_curr(self):
        v = C.c_double()
        self._check(self.L.gpu_get_chi2_curr(self.h, self.nuv, self.npow, self.nbis, self.ndof, self.nuv,
                                             C.byref(v)), "gpu_get_chi2_curr")
        return v.value

    def chi2_noalloc(self, which="ci"):
        """gpair_chi2_noalloc: chi2 of an image buffer with the library's static
        scratch (no device allocation)"""
        out = C.c_double()
        self._check(self.L.gpair_chi2_noalloc(self.h, self.buf(which), C.byref(out)), "gpair_chi2_noalloc")
        return out.value

    def gpu_get_chi2_temp(self):
        v = C.c_double()
        self._check(self.L.gpu_get_chi2_temp(self.h, self.nuv, self.npow, self.nbis, self.ndof, self.nuv,
                                             C.byref(v)), "gpu_get_chi2_temp")
        return v.value

    def gpu_get_entropy_curr(self):
        v = C.c_double()
        self._check(self.L.gpu_get_entropy_curr(self.h, self.W, C.byref(v)), "gpu_get_entropy_curr")
        return v.value

    def gpu_get_entropy_temp(self):
        v = C.c_double()
        self._check(self.L.gpu_get_entropy_temp(self.h, self.W, C.byref(v)), "gpu_get_entropy_temp")
        return v.value

    def gpu_compute_data_gradient_curr(self):
        self._check(self.L.gpu_compute_data_gradient_curr(self.h, self.nuv, self.npow, self.nbis, self.W),
                    "gpu_compute_data_gradient_curr")

    def gpu_compute_data_gradient_temp(self):
        self._check(self.L.gpu_compute_data_gradient_temp(self.h, self.nuv, self.npow, self.nbis, self.W),
                    "gpu_compute_data_gradient_temp")

    def gpu_compute_entropy_gradient_curr(self):
        self._check(self.L.gpu_compute_entropy_gradient_curr(self.h, self.W), "gpu_compute_entropy_gradient_curr")

    def gpu_compute_entropy_gradient_temp(self):
        self._check(self.L.gpu_compute_entropy_gradient_temp(self.h, self.W), "gpu_compute_entropy_gradient_temp")

    def gpu_compute_criterion_gradient_curr(self, alpha):
        self._check(self.L.gpu_compute_criterion_gradient_curr(self.h, self.W, float(alpha)),
                    "gpu_compute_criterion_gradient_curr")

    def gpu_compute_criterion_gradient_temp(self, alpha):
        self._check(self.L.gpu_compute_criterion_gradient_temp(self.h, self.W, float(alpha)),
                    "gpu_compute_criterion_gradient_temp")

    # -- CG ---------------------------------------------------------------------
    def gpu_get_scalprod(self, a, b):
        v = C.c_double()
        self._check(self.L.gpu_get_scalprod(self.h, self.W, self.W, self.buf(a), self.buf(b), C.byref(v)),
                    "gpu_get_scalprod")
        return v.value

    def gpu_compute_descent_dir(self, beta):
        self._check(self.L.gpu_compute_descent_dir(self.h, self.W, float(beta)), "gpu_compute_descent_dir")

    def gpu_update_tempimage(self, step, minval=1e-8, d="dd"):
        self._check(self.L.gpu_update_tempimage(self.h, self.W, float(step), float(minval), self.buf(d)),
                    "gpu_update_tempimage")

    def gpu_update_image(self, step, minval=1e-8, d="dd"):
        self._check(self.L.gpu_update_image(self.h, self.W, float(step), float(minval), self.buf(d)),
                    "gpu_update_image")

    def gpu_backup_gradient(self, src="fgn", dst="fg"):
        self._check(self.L.gpu_backup_gradient(self.h, self.W * self.W, self.buf(src), self.buf(dst)),
                    "gpu_backup_gradient")

    def gpu_linesearch_zoom(self, lo, hi, J_lo, w1, J_init, alpha, max_evals=-1):
        ce, ge, sel = C.c_int(0), C.c_int(0), C.c_double()
        self._check(self.L.gpu_linesearch_zoom(self.h, self.nuv, self.npow, self.nbis, self.ndof, self.nuv, self.W,
                                               float(lo), float(hi), float(J_lo), float(w1), float(J_init),
                                               C.byref(ce), C.byref(ge), self.buf("dd"), self.buf("tg"),
                                               float(alpha), int(max_evals), C.byref(sel)), "gpu_linesearch_zoom")
        return sel.value, ce.value, ge.value

    # -- the loop -----------------------------------------------------------------
    def cg_run(self, alpha, method=2, max_iter=10, chi2r_target=1.0, ls_max_evals=64, verbose=False):
        p = _lib.CgParams()
        self.L.gpair_cg_default_params(C.byref(p))
        p.alpha, p.method, p.max_iter = float(alpha), int(method), int(max_iter)
        p.chi2r_target, p.ls_max_evals, p.verbose = float(chi2r_target), int(ls_max_evals), int(verbose)
        res = _lib.CgResult()
        trace = np.zeros((max(max_iter, 1), 8))
        rc = self.L.gpair_cg_run(self.h, C.byref(p), C.byref(res), _ptr(trace), trace.shape[0])
        self._check(rc, "gpair_cg_run")
        return res, trace[:res.iterations]

    def set_pfactor(self, pf):
        if pf is None:
            self._check(self.L.gpair_set_pfactor(self.h, None, 0), "gpair_set_pfactor")
            return
        a = np.ascontiguousarray(pf, np.float32)
        self._check(self.L.gpair_set_pfactor(self.h, _ptr(a), a.size), "gpair_set_pfactor")

    def autotune(self, reps=5):
        """times candidate launch shapes on this context's problem; returns the chosen
        (line-search chi2 grid, line-search entropy grid, coefficient grid, BF16x3 forward
        rows per group / 32)."""
        out = np.zeros(4, np.int32)
        self._check(self.L.gpair_autotune(self.h, int(reps), _ptr(out), 4), "gpair_autotune")
        return tuple(int(v) for v in out)

    def pcg_run(self, alpha, precond=1, max_iter=1000, chi2r_target=1.0, verbose=False, J_target=0.0):
        """accelerated loop (preconditioned PR-CG with the batched line search);
        J_target > 0: the target also needs an objective at or below it."""
        p = _lib.PcgParams()
        self.L.gpair_pcg_default_params(C.byref(p))
        p.alpha, p.precond, p.max_iter = float(alpha), int(precond), int(max_iter)
        p.chi2r_target, p.verbose, p.J_target = float(chi2r_target), int(verbose), float(J_target)
        res = _lib.PcgResult()
        trace = np.zeros((max(max_iter, 1), 8))
        rc = self.L.gpair_pcg_run(self.h, C.byref(p), C.byref(res), _ptr(trace), trace.shape[0])
        self._check(rc, "gpair_pcg_run")
        return res, trace[:min(res.iterations, trace.shape[0])]

    def pcg_run_checkpointed(self, alpha, precond=1, max_iter=1000, chi2r_target=1.0, resume_from=None,
                             save_to=None, verbose=False):
        """gpair_pcg_run_checkpointed: resume from / save to a checkpoint file"""
        p = _lib.PcgParams()
        self.L.gpair_pcg_default_params(C.byref(p))
        p.alpha, p.precond, p.max_iter = float(alpha), int(precond), int(max_iter)
        p.chi2r_target, p.verbose = float(chi2r_target), int(verbose)
        res = _lib.PcgResult()
        trace = np.zeros((max(max_iter, 1), 8))
        rc = self.L.gpair_pcg_run_checkpointed(self.h, C.byref(p), C.byref(res), _ptr(trace), trace.shape[0],
                                               None if resume_from is None else str(resume_from).encode(),
                                               None if save_to is None else str(save_to).encode())
        self._check(rc, "gpair_pcg_run_checkpointed")
        return res, trace[:min(res.iterations, trace.shape[0])]

    def set_snapshot(self, path, every=5, pixsize=0.0):
        """the loops write the current image to `path` every `every` iterations (None: off)"""
        self._check(self.L.gpair_set_snapshot(self.h, None if path is None else str(path).encode(), int(every),
                                              float(pixsize)), "gpair_set_snapshot")

    def set_sparse_commit(self, max_pixels=-1):
        """gpair_set_sparse_commit: -1 default cap (W^2/40), 0 always the full DFT"""
        self._check(self.L.gpair_set_sparse_commit(self.h, int(max_pixels)), "gpair_set_sparse_commit")

    def set_adjoint_tile(self, rows, cols, workgroups=0):
        """gpair_set_adjoint_tile: 256 x 256, 128 x 256 or 128 x 128; about `workgroups` in total (0: 256)"""
        self._check(self.L.gpair_set_adjoint_tile(self.h, int(rows), int(cols), int(workgroups)),
                    "gpair_set_adjoint_tile")

    def set_graphs(self, on=True):
        self._check(self.L.gpair_set_graphs(self.h, int(bool(on))), "gpair_set_graphs")


def batch_pcg_run(ctxs, alphas, precond=1, max_iter=1000, chi2r_target=1.0, streams=False):
    """gpair_batch_pcg_run over several contexts (one launch sequence per iteration
    for all of them where they qualify; streams=True: gpair_batch_pcg_run_streams,
    one stream each); returns the per-context results."""
    L = _lib.load()
    n = len(ctxs)
    params = (_lib.PcgParams * n)()
    for i, a in enumerate(alphas):
        L.gpair_pcg_default_params(C.byref(params[i]))
        params[i].alpha, params[i].precond, params[i].max_iter = float(a), int(precond), int(max_iter)
        params[i].chi2r_target = float(chi2r_target)
    results = (_lib.PcgResult * n)()
    handles = (C.c_void_p * n)(*[c.h.value for c in ctxs])
    rc = (L.gpair_batch_pcg_run_streams if streams else L.gpair_batch_pcg_run)(handles, n, params, results)
    if rc != 0:
        bad = [i for i in range(n) if results[i].status]
        raise GpairError(f"gpair_batch_pcg_run: {L.gpair_strerror(rc).decode()} (contexts {bad})")
    return list(results)
