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


GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("name", ["ref_cpu_w64.npz", "ref_cpu_w64_7tel.npz"])
def test_set_model_matches_reference_priors(name):
    """gpair_set_model types 0-4 against the reference's set_model output
    (gpair1/trunk/cpu.c:645-749, width 1 mas, flux 1, the reference's own offsets)."""
    z = np.load(os.path.join(GOLDEN_DIR, name), allow_pickle=False)
    W, pix = int(z["W"]), float(z["pix"])
    for t in range(5):
        got = synth.set_model(W, pix, t, 1.0, 1.0, centered=False)
        ref = z["ref_priors"][t]
        assert np.array_equal(got, ref), (t, np.abs(got - ref).max())
    with pytest.raises(ValueError):
        synth.set_model(W, pix, 7, 1.0, 1.0)


def _init_data(pow_, powerr, amp, amperr, phs, phserr, uvpnt, extrap=1):
    L = _lib.load()
    npow, nbis = pow_.size, amp.size
    with np.errstate(over="ignore"):        # 1e99 is +inf in float, as in the reference
        amp, amperr = amp.astype(np.float32).copy(), amperr.astype(np.float32).copy()
    data = np.zeros(npow + 2 * nbis, np.float32)
    err = np.zeros_like(data)
    ph = np.zeros(2 * nbis, np.float32)
    nd = C.c_int()
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    args = [f32(pow_), f32(powerr), amp, amperr, f32(phs), f32(phserr), np.ascontiguousarray(uvpnt, np.int64)]
    rc = L.gpair_init_data(npow, nbis, *[a.ctypes.data for a in args], extrap, data.ctypes.data, err.ctypes.data,
                           ph.ctypes.data, C.byref(nd))
    assert rc == 0
    return data, err, ph, nd.value, amp, amperr


def test_init_data_closed_forms():
    """gpair_init_data (gpair1/trunk/gpair.c:997-1066): packing, and the T3
    amplitude extrapolation from the three V2 points checked through the
    quadratic its estimate solves and its small-error limit |V_ab||V_bc||V_ca|."""
    r = np.random.default_rng(1)
    npow, nbis = 12, 5
    pow_ = r.uniform(0.05, 0.9, npow)
    powerr = 0.02 * pow_
    amp = r.uniform(0.01, 0.5, nbis)
    amperr = 0.05 * amp
    phs = r.uniform(-180, 180, nbis)
    phserr = np.full(nbis, 1.0)
    uvpnt = r.integers(0, npow, size=(nbis, 3))
    amperr[1] = -1.0          # missing amplitude: extrapolated
    amperr[3] = 1e99          # NaN amplitudes arrive as 1 +- 1e99
    uvpnt[4, 2] = npow + 3    # a leg without V2: no amplitude weight, one dof less
    amperr[4] = 0.0
    data, err, ph, nd, a2, e2 = _init_data(pow_, powerr, amp, amperr, phs, phserr, uvpnt)
    p32, e32 = pow_.astype(np.float32), powerr.astype(np.float32)
    assert np.array_equal(data[:npow], p32) and np.allclose(err[:npow], 1 / e32, rtol=1e-6)
    assert np.all(data[npow + 1::2] == 0)
    assert np.allclose(ph[0::2] + 1j * ph[1::2], np.exp(-1j * phs.astype(np.float32) * np.pi / 180), atol=1e-6)
    assert nd == npow + 2 * nbis - 1
    for t in (0, 2):     # measured amplitudes pass through
        assert data[npow + 2 * t] == np.float32(amp[t]) and np.isclose(err[npow + 2 * t], 1 / amperr[t], rtol=1e-6)
    for t in (1, 3):
        s = []
        for k in uvpnt[t]:
            p, e = float(p32[k]), float(e32[k])
            st = (p + np.sqrt(p * p + 2 * e * e)) / 2
            assert abs(st * st - p * st - e * e / 2) <= 1e-6 * st * st       # the quadratic
            s.append(st)
        assert np.isclose(a2[t], np.sqrt(np.prod(s)), rtol=1e-5)
        assert data[npow + 2 * t] == a2[t] and e2[t] > 0
    assert a2[4] == 1.0 and err[npow + 8] == 0.0


def test_init_data_small_error_limit():
    """small V2 errors: the extrapolated amplitude tends to |V_ab||V_bc||V_ca|"""
    pow_ = np.array([0.25, 0.36, 0.49, 0.64])
    uvpnt = np.array([[0, 1, 2], [1, 2, 3]])
    data, err, ph, nd, a2, e2 = _init_data(pow_, 1e-6 * pow_, np.ones(2), -np.ones(2), np.zeros(2), np.ones(2),
                                           uvpnt)
    assert np.allclose(a2, [0.5 * 0.6 * 0.7, 0.6 * 0.7 * 0.8], rtol=1e-5)
    # relative error of the product of three independent 1e-6 estimates ~ sqrt(3) * 1e-6 / 2
    assert np.allclose(e2 / a2, np.sqrt(3) * 0.5e-6, rtol=0.05)


def test_batch_rejects_bad_arguments():
    L = _lib.load()
    p = (_lib.PcgParams * 2)()
    r = (_lib.PcgResult * 2)()
    assert L.gpair_batch_pcg_run(None, 2, p, r) == 1
    h = (C.c_void_p * 2)(None, None)
    assert L.gpair_batch_pcg_run(h, 2, p, r) == 1
    assert L.gpair_batch_pcg_run(h, 0, p, r) == 1
    assert L.gpair_set_graphs(None, 1) == 1
