"""ctypes binding of libgpair_hip.so (include/gpair_hip.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpair_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "gpair_hip.h")
HEADERS = [HEADER, os.path.join(os.path.dirname(HERE), "include", "gpair_nn.h"),
           os.path.join(os.path.dirname(HERE), "include", "gpair_io.h")]

_vp = C.c_void_p
_i = C.c_int
_d = C.c_double
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_l = C.c_long
_f = C.c_float


class CgParams(C.Structure):
    _fields_ = [("alpha", _d), ("method", _i), ("max_iter", _i), ("chi2r_target", _d), ("ls_max_evals", _i),
                ("minvalue", _d), ("wolfe1", _d), ("wolfe2", _d), ("step_max", _d), ("verbose", _i)]


class CgResult(C.Structure):
    _fields_ = [("iterations", _i), ("criterion_evals", _i), ("grad_evals", _i), ("J", _d), ("chi2", _d),
                ("entropy", _d), ("chi2r", _d), ("converged", _i), ("status", _i)]


class PcgParams(C.Structure):
    _fields_ = [("alpha", _d), ("precond", _i), ("max_iter", _i), ("chi2r_target", _d), ("minvalue", _d),
                ("wolfe1", _d), ("verbose", _i), ("J_target", _d)]


class PcgResult(C.Structure):
    _fields_ = [("iterations", _i), ("forward_dfts", _i), ("adjoint_dfts", _i), ("ls_batches", _i), ("J", _d),
                ("chi2", _d), ("entropy", _d), ("chi2r", _d), ("converged", _i), ("stalled", _i), ("status", _i),
                ("resyncs", _i), ("sparse_commits", _i)]


# name -> (restype, argtypes)
class OiData(C.Structure):
    """gpair_oidata (include/gpair_io.h)"""
    _fields_ = [("npow", C.c_int), ("nbis", C.c_int), ("nuv", C.c_int),
                ("pow", C.POINTER(C.c_float)), ("powerr", C.POINTER(C.c_float)),
                ("bisamp", C.POINTER(C.c_float)), ("bisamperr", C.POINTER(C.c_float)),
                ("bisphs", C.POINTER(C.c_float)), ("bisphserr", C.POINTER(C.c_float)),
                ("u", C.POINTER(C.c_float)), ("v", C.POINTER(C.c_float)),
                ("wavelength", C.POINTER(C.c_float)), ("bandwidth", C.POINTER(C.c_float)),
                ("uvpnt", C.POINTER(C.c_int64)), ("sign", C.POINTER(C.c_int16)),
                ("target_id", C.c_int), ("minband", C.c_float), ("maxband", C.c_float),
                ("uv_compares", C.c_longlong)]


SIGNATURES = {
    "gpair_strerror": (C.c_char_p, [_i]),
    "gpair_last_error": (C.c_char_p, [_vp]),
    "gpu_init": (_i, [C.POINTER(_vp), _i]),
    "gpu_copy_data": (_i, [_vp, _vp, _vp, _i, _i, _vp, _i, _i, _vp, _vp, _i, _vp, _i, _i]),
    "gpu_copy_image": (_i, [_vp, _vp, _i, _i]),
    "gpu_copy_dft": (_i, [_vp, _vp, _vp, _i]),
    "gpair_set_uv": (_i, [_vp, _vp, _vp, _i, C.c_float]),
    "gpu_build_kernels": (_i, [_vp, _i, _i, _i]),
    "gpu_cleanup": (_i, [_vp]),
    "gpu_shutdown": (_i, [_vp]),
    "gpair_set_precision": (_i, [_vp, _i]),
    "gpu_getp_ci": (_vp, [_vp]),
    "gpu_getp_ti": (_vp, [_vp]),
    "gpu_getp_fgn": (_vp, [_vp]),
    "gpu_getp_fg": (_vp, [_vp]),
    "gpu_getp_dd": (_vp, [_vp]),
    "gpu_getp_tg": (_vp, [_vp]),
    "gpu_getp_eg": (_vp, [_vp]),
    "gpair_getp_dg": (_vp, [_vp]),
    "gpu_get_chi2_curr": (_i, [_vp, _i, _i, _i, _i, _i, _pd]),
    "gpu_get_chi2_temp": (_i, [_vp, _i, _i, _i, _i, _i, _pd]),
    "gpu_get_entropy_curr": (_i, [_vp, _i, _pd]),
    "gpu_get_entropy_temp": (_i, [_vp, _i, _pd]),
    "gpu_compute_data_gradient_curr": (_i, [_vp, _i, _i, _i, _i]),
    "gpu_compute_data_gradient_temp": (_i, [_vp, _i, _i, _i, _i]),
    "gpu_compute_entropy_gradient_curr": (_i, [_vp, _i]),
    "gpu_compute_entropy_gradient_temp": (_i, [_vp, _i]),
    "gpu_compute_criterion_gradient_curr": (_i, [_vp, _i, _d]),
    "gpu_compute_criterion_gradient_temp": (_i, [_vp, _i, _d]),
    "gpu_get_scalprod": (_i, [_vp, _i, _i, _vp, _vp, _pd]),
    "gpu_compute_descent_dir": (_i, [_vp, _i, _d]),
    "gpu_update_tempimage": (_i, [_vp, _i, _d, _d, _vp]),
    "gpu_update_image": (_i, [_vp, _i, _d, _d, _vp]),
    "gpu_backup_gradient": (_i, [_vp, _i, _vp, _vp]),
    "gpu_linesearch_zoom": (_i, [_vp, _i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _d, _pi, _pi, _vp, _vp, _d, _i, _pd]),
    "gpu_get_image": (_i, [_vp, _i, _vp, _vp]),
    "gpair_get_vis": (_i, [_vp, _vp, _i]),
    "gpair_get_mock": (_i, [_vp, _vp, _i]),
    "gpair_put_image": (_i, [_vp, _vp, _vp, _i]),
    "gpair_cg_default_params": (None, [C.POINTER(CgParams)]),
    "gpair_cg_run": (_i, [_vp, C.POINTER(CgParams), C.POINTER(CgResult), _vp, _i]),
    "gpair_sizes": (_i, [_vp, _pi, _pi, _pi, _pi]),
    "gpair_sync": (_i, [_vp]),
    "gpair_timing_enable": (_i, [_vp, _i]),
    "gpair_timing_read": (_i, [_vp, _i, _pi, _pd]),
    "gpair_pcg_default_params": (None, [C.POINTER(PcgParams)]),
    "gpair_pcg_run": (_i, [_vp, C.POINTER(PcgParams), C.POINTER(PcgResult), _vp, _i]),
    "gpair_set_pfactor": (_i, [_vp, _vp, _i]),
    "gpair_autotune": (_i, [_vp, _i, _vp, _i]),
    "gpair_set_graphs": (_i, [_vp, _i]),
    "gpair_set_sparse_commit": (_i, [_vp, _i]),
    "gpair_set_adjoint_tile": (_i, [_vp, _i, _i, _i]),
    "gpair_pcg_run_checkpointed": (_i, [_vp, C.POINTER(PcgParams), C.POINTER(PcgResult), _vp, _i, C.c_char_p,
                                        C.c_char_p]),
    "gpair_set_snapshot": (_i, [_vp, C.c_char_p, _i, _f]),
    "gpair_batch_pcg_run": (_i, [_vp, _i, _vp, _vp]),
    "gpair_batch_pcg_run_streams": (_i, [_vp, _i, _vp, _vp]),
    "gpair_pixel_begin": (_i, [_vp]),
    "gpu_update_vis_fluxchange": (_i, [_vp, _i, _i, _f, _i, _i, _i]),
    "gpair_update_vis_positionchange": (_i, [_vp, _i, _i, _i, _i, _f, _i, _i]),
    "gpu_new_chi2": (_i, [_vp, _i, _i, _i, _i, _pd]),
    "gpair_get_trial_vis": (_i, [_vp, _vp, _i]),
    "gpair_pixel_propose": (_i, [_vp, _vp, _i, _vp, _vp]),
    "gpair_pixel_apply": (_i, [_vp, _vp, _i]),
    "gpair_pixel_state": (_i, [_vp, _pd, _pd, _pd]),
    "gpair_chi2_noalloc": (_i, [_vp, _vp, _pd]),
    # data I/O (include/gpair_io.h)
    "gpair_io_last_error": (C.c_char_p, []),
    "gpair_oifits_targets": (_i, [C.c_char_p, _vp, _vp, _i, _pi]),
    "gpair_oifits_read": (_i, [C.c_char_p, _i, _f, _f, C.POINTER(OiData)]),
    "gpair_oidata_free": (None, [C.POINTER(OiData)]),
    "gpair_oifits_write": (_i, [C.c_char_p, C.c_char_p, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _i,
                                _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gpair_fits_read_image": (_i, [C.c_char_p, _vp, C.c_long, _pi, _pi]),
    "gpair_fits_write_image": (_i, [C.c_char_p, _vp, _i, _i, _i, _f, C.c_char_p, C.c_char_p]),
    "gpair_set_model": (_i, [_vp, _i, C.c_float, _i, C.c_float, C.c_float, _i]),
    "gpair_init_data": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _pi]),
    # learned start (include/gpair_nn.h)
    "gpair_nn_alloc": (_vp, [_l]),
    "gpair_nn_free": (_i, [_vp]),
    "gpair_nn_upload": (_i, [_vp, _vp, _l]),
    "gpair_nn_download": (_i, [_vp, _vp, _l]),
    "gpair_nn_conv3_fwd": (_i, [_vp, _l, _i, _vp, _vp, _vp, _l, _i, _i, _i, _i, _i]),
    "gpair_nn_conv3_bwd": (_i, [_vp, _l, _i, _vp, _vp, _l, _i, _i, _i, _i, _vp, _l, _vp, _vp]),
    "gpair_nn_pool2": (_i, [_vp, _l, _vp, _l, _i, _i, _i, _i, _i]),
    "gpair_nn_up2": (_i, [_vp, _l, _vp, _l, _i, _i, _i, _i, _i]),
    "gpair_unet_create": (_i, [C.POINTER(_vp), _i, _i, _i, _i, _i, _i, C.c_ulonglong]),
    "gpair_unet_destroy": (_i, [_vp]),
    "gpair_unet_num_params": (_l, [_vp]),
    "gpair_unet_get_params": (_i, [_vp, _vp, _l]),
    "gpair_unet_set_params": (_i, [_vp, _vp, _l]),
    "gpair_unet_get_grads": (_i, [_vp, _vp, _l]),
    "gpair_unet_forward": (_i, [_vp, _vp, _i, _vp]),
    "gpair_unet_backward": (_i, [_vp, _vp, _i]),
    "gpair_unet_adam": (_i, [_vp, _f]),
    "gpair_unet_train_step": (_i, [_vp, _vp, _vp, _i, _f, _pd]),
}

_LIB = None


def load(path: str = LIB_PATH):
    """Load and type the library; raises when it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m gpair_amd.build` (there is no CPU fallback)")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def header_symbols(paths=None):
    """function names declared in include/gpair_hip.h and include/gpair_nn.h"""
    import re
    txt = "\n".join(open(p).read() for p in (paths or HEADERS))
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt)
    skip = {"if", "defined"}
    return sorted({n for n in names if n.startswith(("gpu_", "gpair_")) and n not in skip})
