"""CPU-side checks: the C ABI library, the synthetic problem packing, bench contract."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

from gpair_amd import _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_header_symbol():
    L = _lib.load()
    names = _lib.header_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # every declared function has a ctypes signature, and vice versa
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_error_strings_and_null_handling():
    L = _lib.load()
    assert L.gpair_strerror(0) == b"ok"
    assert L.gpair_strerror(1) == b"invalid argument"
    # NULL context/handles are rejected without touching a device
    assert L.gpu_cleanup(None) == 1
    assert L.gpair_sync(None) == 1
    assert L.gpu_init(None, 0) == 1
    assert L.gpu_getp_ci(None) is None
    p = _lib.CgParams()
    L.gpair_cg_default_params(C.byref(p))
    assert p.method == 2 and p.minvalue == 1e-8 and p.wolfe1 == 1e-4 and p.wolfe2 == 0.1 and p.step_max == 100.0
    assert L.gpair_cg_run(None, C.byref(p), None, None, 0) == 1


def test_init_data_packing():
    """synth packs data exactly as init_data does (gpair1/trunk/gpair.c:1000-1063)."""
    prob = synth.make_problem("cfg1", seed=3)
    n = prob.npow
    assert prob.ndof == n + 2 * prob.nbis
    assert np.all(prob.data[n + 1::2] == 0.0)
    assert np.all(prob.err > 0)
    assert np.allclose(np.abs(prob.phasor), 1.0, atol=1e-6)
    # V2 errors are 2 % relative with a 1e-4 floor; T3 amplitude errors 5 % relative
    v2_true = np.abs(prob.truth_vis) ** 2
    assert np.allclose(1.0 / prob.err[:n], np.maximum(0.02 * v2_true, 1e-4), rtol=1e-5)


def test_layout_closure_and_signs():
    """every T3 leg points at a stored V2 point; with the signs the legs close (u_ab + u_bc + u_ca = 0)."""
    prob = synth.make_problem("cfg1", seed=4)
    assert prob.npow == prob.nuv                      # all legs are V2 baselines
    assert prob.nbis == prob.meta["ntri"] * prob.meta["nch"] * prob.meta["nsnap"]
    assert np.all(prob.u >= 0)                        # +u half plane (getoifits.c:147-151)
    uv = np.stack([prob.u, prob.v], 1).astype(np.float64)
    legs = uv[prob.uvpnt] * prob.sign[..., None]
    closure = np.abs(legs.sum(axis=1)).max()
    assert closure < 1e-6 * np.abs(uv).max()
    assert set(np.unique(prob.sign)) <= {-1, 1}


def test_configs_sizes():
    """the three BASELINE configurations have the stated sizes."""
    for cfg, (W, npow, nbis) in {"cfg1": (64, 480, 640), "cfg2": (256, 36864, 86016)}.items():
        c = synth.CONFIGS[cfg]
        rng = np.random.default_rng(0)
        uv, uvpnt, sign, (nbase, ntri) = synth.build_layout(rng, c["ntel"], c["nch"], c["nsnap"], c["uvmax"])
        assert c["W"] == W and uv.shape[0] == npow and uvpnt.shape[0] == nbis
    c = synth.CONFIGS["cfg3"]
    assert c["W"] == 512 and c["ntel"] * (c["ntel"] - 1) // 2 * c["nch"] * c["nsnap"] == 147456


def test_truth_image_fits_its_data():
    """the grid-rendered sky reproduces its own noiseless visibilities"""
    prob = synth.make_problem("cfg1", seed=2)
    au, av = prob.au(), prob.av()
    V = synth.dft64(prob.truth.astype(np.float64), prob.W, au, av)
    assert np.abs(V - prob.truth_vis).max() < 1e-6


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stdout
