"""Benchmark of the MEM-CG reconstruction loop on one MI355X.

    python bench.py [--steps K] [--warmup W] [--cfg cfg2] [--to-target] [--no-cpu]

A step is one CG iteration of the reference's GPU-path loop (objective,
gradient, beta, descent direction, bracket + zoom line search, image update),
driven by the C host loop gpair_cg_run over the C ABI, on the 256x256 headline
problem (36,864 V2 + 86,016 T3, synthetic, seeded).  Inputs are resident in HBM
before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("wall-time & CG iters/s to chi2_r=1 at 256², BF16x3/learned vs plain FP32; err vs FP64")
PEAK_FP32_MFMA = 157.3   # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0        # GB/s
PROFILES = os.path.join(ROOT, "profiles")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cfg", default="cfg2")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--alpha", type=float, default=None)
    ap.add_argument("--method", type=int, default=2)
    ap.add_argument("--ls-max-evals", type=int, default=64)
    ap.add_argument("--to-target", action="store_true", help="also time the run to chi2_r = 1")
    ap.add_argument("--max-iter", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-kind", default="auto", choices=["auto", "reference", "port"])
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 256 MiB device copy (HBM counter calibration for rocprofv3 --pmc)")
    return ap.parse_args()


def calib_copy():
    """one device-to-device copy of 256 MiB: the FETCH_SIZE / WRITE_SIZE counters of
    its kernel give the KiB-per-byte scale used to turn counters into HBM bytes."""
    import torch
    n = 64 * 2 ** 20
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calib_copy_bytes": n * 4}))


def read_traffic(W):
    """HBM bytes per forward-DFT launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(PROFILES, f"forward_dft_W{W}_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def cpu_baseline(prob, evals_J, evals_G, kind):
    """Reference CPU model (oracle/_ref, 1 core) or the FP64 port, on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if kind == "auto":
        kind = "reference" if O.ref_available() else "port"
    if kind == "reference":
        from gpair_amd import synth
        R = O.RefCPU(prob)
        img = prob.start.copy()
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        sys.stdout.flush()
        os.dup2(devnull, 1)        # the reference prints "CPU Flux" per call
        try:
            t0 = time.perf_counter()
            R.image2chi2(img)
            tJ = time.perf_counter() - t0
            # gradient: same uv set on a 64x64 grid, cost is linear in pixel count
            small = synth.make_problem(prob.meta["cfg"], seed=prob.meta["seed"], W=64,
                                       pix=prob.pixellation * prob.W / 64)
            Rs = O.RefCPU(small)
            Rs.image2chi2(small.start)
            t0 = time.perf_counter()
            Rs.data_gradient(small.start)
            tG = (time.perf_counter() - t0) * (prob.W * prob.W) / (64 * 64)
        finally:
            os.dup2(saved, 1)
            os.close(devnull)
        per_iter = evals_J * tJ + evals_G * tG
        sample = (f"reference cpu.c (image2chi2 at {prob.W}^2 = {tJ:.2f} s; compute_data_gradient on the same uv "
                  f"at 64^2 x{prob.W * prob.W // 4096} = {tG:.1f} s) x ({evals_J:.1f} J + {evals_G:.1f} grad) "
                  f"evals per CG iteration from the GPU run")
        return {"value": 1.0 / per_iter, "unit": "CG iter/s", "cores": 1, "kind": "reference", "sample": sample}
    orc = O.Oracle()
    P = orc.unpack(prob)
    img = prob.start.astype(np.float64)
    t0 = time.perf_counter()
    J = orc.objective(P, img, 100.0)
    tJ = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.gradient(P, img, 100.0, J[3], J[4])
    tG = time.perf_counter() - t0
    per_iter = evals_J * tJ + evals_G * tG
    return {"value": 1.0 / per_iter, "unit": "CG iter/s", "cores": orc.threads(), "kind": "port",
            "sample": f"FP64 oracle: one J ({tJ:.2f} s) + one gradient ({tG:.2f} s) scaled by the evals per iteration"}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    if a.calib_copy:
        calib_copy()
        return
    from gpair_amd import _lib, synth
    from gpair_amd.api import GpairContext

    prob = synth.make_problem(a.cfg, seed=a.seed)
    alpha = a.alpha if a.alpha is not None else prob.meta["alpha"]
    print(f"# problem {a.cfg}: W={prob.W} nuv={prob.nuv} nbis={prob.nbis} ndof={prob.ndof} alpha={alpha}",
          file=sys.stderr, flush=True)
    ctx = GpairContext.from_problem(prob)
    L = ctx.L
    if a.warmup > 0:
        ctx.cg_run(alpha, method=a.method, max_iter=a.warmup, chi2r_target=0.0, ls_max_evals=a.ls_max_evals)
    L.gpair_sync(ctx.h)
    L.gpair_timing_enable(ctx.h, 1)
    t0 = time.perf_counter()
    res, tr = ctx.cg_run(alpha, method=a.method, max_iter=a.steps, chi2r_target=0.0, ls_max_evals=a.ls_max_evals)
    L.gpair_sync(ctx.h)
    dt = time.perf_counter() - t0
    cnt, tot = C.c_int(), C.c_double()
    L.gpair_timing_read(ctx.h, 0, C.byref(cnt), C.byref(tot))
    fwd_n, fwd_ms = cnt.value, tot.value
    L.gpair_timing_read(ctx.h, 1, C.byref(cnt), C.byref(tot))
    adj_n, adj_ms = cnt.value, tot.value
    L.gpair_timing_enable(ctx.h, 0)
    print(f"# timed {a.steps} CG iterations in {dt:.3f} s: J evals {res.criterion_evals}, grad evals "
          f"{res.grad_evals}, chi2r {res.chi2r:.4f}", file=sys.stderr, flush=True)

    W, nuv = prob.W, prob.nuv
    flop_fwd = 4.0 * nuv * W * W + 8.0 * nuv * W
    avg_fwd = fwd_ms / max(fwd_n, 1) * 1e-3
    achieved = flop_fwd / avg_fwd / 1e12
    avg_adj = adj_ms / max(adj_n, 1) * 1e-3
    out = {
        "metric": METRIC,
        "value": a.steps / dt,
        "unit": "CG iter/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded array + analytic sky, noise per BASELINE configs)",
        "config": {"workload": f"{a.cfg}: plain MEM-CG, FP32 MFMA DFT, FP64 scalars; {W}x{W} px, "
                               f"{prob.npow} V2 + {prob.nbis} T3",
                   "W": W, "nuv": nuv, "npow": prob.npow, "nbis": prob.nbis, "alpha": alpha,
                   "method": ["SD", "FR", "PR", "HS"][a.method], "seed": a.seed, "precision": "fp32"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_MFMA, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_MFMA, "traffic": read_traffic(W),
                     "kernel": "k_vis_fwd_f32 (forward DFT)", "avg_us": avg_fwd * 1e6, "launches": fwd_n,
                     "flop_per_launch": flop_fwd},
        "adjoint": {"kernel": "k_adj_f32", "avg_us": avg_adj * 1e6, "launches": adj_n,
                    "achieved_tflops": flop_fwd / avg_adj / 1e12 if adj_n else None},
        "evals_per_iter": {"J": res.criterion_evals / a.steps, "grad": res.grad_evals / a.steps},
        "chi2r_end": res.chi2r,
    }
    if a.to_target:
        ctx2 = GpairContext.from_problem(prob)
        t0 = time.perf_counter()
        r2, tr2 = ctx2.cg_run(alpha, method=a.method, max_iter=a.max_iter, chi2r_target=1.0,
                              ls_max_evals=a.ls_max_evals)
        ctx2.L.gpair_sync(ctx2.h)
        out["to_target"] = {"wall_s": time.perf_counter() - t0, "iterations": r2.iterations,
                            "converged": bool(r2.converged), "chi2r": r2.chi2r, "J": r2.J}
        ctx2.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, res.criterion_evals / a.steps, res.grad_evals / a.steps,
                                           a.cpu_kind)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
