"""Benchmark of the MEM-CG reconstruction loop on one MI355X.

    python bench.py [--steps K] [--warmup W] [--cfg cfg2] [--no-cpu] [--plain-budget-s S]

Workload: the 256x256 headline problem (36,864 V2 + 86,016 T3, synthetic and
seeded, Gaussian prior of FWHM 5 mas), all inputs resident in HBM before any
timed region.  A step is one CG iteration.

* value: CG iterations per second of the accelerated loop (gpair_pcg_run:
  image-metric preconditioned PR-CG, batched line search with FP64 objective
  changes) over K timed iterations after W warm-up iterations.
* to_target: wall time to chi2/ndof <= 1 from the prior for the accelerated
  loop in FP32, in BF16x3 and with FP16 DFTs (power-of-two scaling; convergence
  confirmed on an exact BF16x3 evaluation), and for the plain loop with the
  reference's rules (gpair_cg_run, FP32 DFT), the latter bounded by
  --plain-budget-s.
* learned_cfg5: the learned start against the plain start on held-out problems
  of the network's training family (256x256, 6-telescope array), network
  inference included, BF16x3.
* roofline: the forward DFT, timed with HIP events on the context stream, against
  the roofline that binds it.  FP32 reads the full X and Y tables (64 flop per
  compulsory byte at 256^2, above the FP32 balance of 20): MFMA, against 157.3
  TFLOP/s.  BF16x3 rebuilds table entries from the factored tables (8 (W/16 + 16)
  bytes per uv point and axis: 502 flop per compulsory byte at 256^2, above the
  BF16 balance of 312): MFMA.  frac = algorithmic flops per launch / launch time /
  the published dense BF16 (or FP16) peak of 2.5 PFLOP/s; mfma_pipe_frac counts the
  MFMA work actually issued (3 BF16 products per multiply-add) against the same peak.
* cpu_baseline: the reference's own CPU model (oracle/_ref) on one core; the
  same leg re-evaluates each final image's objective with the FP64 oracle
  (err_vs_fp64).
* batch_cfg4: many independent reconstructions of uneven sizes, sequential
  against gpair_batch_pcg_run (one launch sequence per iteration for all
  problems) and gpair_batch_pcg_run_streams (one stream per problem).
* precision_cfg3: the 512x512 precision variant, FP32 / BF16x3 / FP16 with scaling
  against the FP64 oracle (visibilities, chi2, gradient, DFT times, and the
  objective after a bounded run).
* pixel_moves: single-pixel moves scored per second (gpair_pixel_propose, --moves
  per call) against a full forward DFT + chi2 per move, with the kernels' own
  rooflines and the FP64 port on the host cores.
Prints one JSON line on stdout.
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

METRIC = "wall-time & CG iters/s to chi2_r=1 at 256², BF16x3/learned vs plain FP32; err vs FP64"
PEAK = {"fp32": 157.3, "bf16x3": 2500.0, "fp16s": 2500.0}   # TFLOP/s, dense MFMA: FP32, BF16, FP16
PRODUCTS = {"fp32": 1, "bf16x3": 3, "fp16s": 1}               # MFMA products per algorithmic multiply-add
HBM_PEAK = 8.0e12                                  # B/s
TIMING_EVERY = 8   # DFT launches per timed launch in the headline loop
PROFILES = os.path.join(ROOT, "profiles")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cfg", default="cfg2")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--alpha", type=float, default=None)
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16x3", "fp16s"])
    ap.add_argument("--max-iter", type=int, default=30000)
    ap.add_argument("--plain-budget-s", type=float, default=30.0)
    ap.add_argument("--no-target", action="store_true", help="skip the runs to chi2_r = 1")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-start workload")
    ap.add_argument("--learned-problems", type=int, default=128, help="held-out problems of the learned-start leg")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in launch shapes")
    ap.add_argument("--batch", type=int, default=64, help="problems of the batching workload (0: skip)")
    ap.add_argument("--spread-seeds", default="2,3,4",
                    help="further seeds of the headline configuration run to the target with BF16x3 "
                         "(the iteration count to chi2_r = 1 varies with the trajectory)")
    ap.add_argument("--batch-iters", type=int, default=100)
    ap.add_argument("--batch-modes", default="streams,grouped",
                    help="batched runs after the sequential one: streams (one stream per problem), grouped "
                         "(one launch sequence per iteration)")
    ap.add_argument("--moves", type=int, default=4096, help="single-pixel moves scored per call (0: skip)")
    ap.add_argument("--no-cfg3", action="store_true", help="skip the 512x512 precision workload")
    ap.add_argument("--cfg3-iters", type=int, default=300, help="accelerated iterations per precision at 512x512")
    ap.add_argument("--move-calls", type=int, default=20, help="timed gpair_pixel_propose calls")
    ap.add_argument("--cpu-kind", default="auto", choices=["auto", "reference", "port"])
    ap.add_argument("--ref-loop-s", type=float, default=20.0,
                    help="seconds of the reference's own CPU program on the 64x64 workload (0: skip)")
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


def read_traffic(name, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a device copy of known size), if present."""
    path = os.path.join(PROFILES, name)
    if not os.path.exists(path):
        return None
    seen = False
    for line in open(path):
        if line.startswith(kernel + ":"):
            seen = "not dispatched" not in line
        elif seen and line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
        elif not line.strip():
            seen = False   # each kernel's figures end at a blank line
    return None


def cpu_baseline(prob, evals_J, evals_G, kind):
    """The reference's CPU model (oracle/_ref, one core) or the FP64 port, on a bounded
    sample scaled to CG iterations of the reference rules."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if kind == "auto":
        kind = "reference" if O.ref_available() else "port"
    if kind == "reference":
        from gpair_amd import synth
        R = O.RefCPU(prob)
        devnull = os.open(os.devnull, os.O_WRONLY)
        saved = os.dup(1)
        sys.stdout.flush()
        os.dup2(devnull, 1)        # the reference prints a line per image2vis call
        try:
            t0 = time.perf_counter()
            R.image2chi2(prob.start)
            tJ = time.perf_counter() - t0
            # gradient: the same uv set on a 64x64 grid; its cost is linear in the pixel count
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
        sample = (f"reference cpu.c: image2chi2 at {prob.W}^2 {tJ:.2f} s, compute_data_gradient on the same uv "
                  f"at 64^2 x{prob.W * prob.W // 4096} = {tG:.1f} s; per CG iteration of the reference rules "
                  f"{evals_J:.1f} J + {evals_G:.1f} gradient evaluations (counts from the GPU run)")
        return {"value": 1.0 / per_iter, "unit": "CG iter/s", "cores": 1, "kind": "reference", "sample": sample}
    orc = O.Oracle()
    P = orc.unpack(prob)
    img = prob.start.astype(np.float64)
    t0 = time.perf_counter()
    J = orc.objective(P, img, 1.0)
    tJ = time.perf_counter() - t0
    t0 = time.perf_counter()
    orc.gradient(P, img, 1.0, J[3], J[4])
    tG = time.perf_counter() - t0
    per_iter = evals_J * tJ + evals_G * tG
    return {"value": 1.0 / per_iter, "unit": "CG iter/s", "cores": orc.threads(), "kind": "port",
            "sample": f"FP64 oracle: one J ({tJ:.2f} s) + one gradient ({tG:.2f} s) x evals per iteration"}


def reference_loop_cfg1(a):
    """The reference's own CPU reconstruction program (gpair1/branches/cpu_dev, built
    by oracle/Makefile into oracle/_ref/gpair_cpu) on the 64x64 workload (480 V2 +
    640 T3, flat prior, alpha 0.025) written as OIFITS, pinned to one core and stopped
    after a.ref_loop_s seconds; its CG iterations are counted from the per-iteration
    line it prints.  Beside it, the same OIFITS file through the library: the
    reference rules (gpair_cg_run, FP32 DFTs) for as many iterations, and the
    accelerated loop to chi2_r = 1."""
    import subprocess
    import tempfile
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from gpair_amd import oifits, synth
    from gpair_amd.api import GpairContext
    if not os.path.exists(O.REF_LOOP):
        return None
    prob = synth.make_problem("cfg1", seed=a.seed)
    W, alpha = prob.W, prob.meta["alpha"]
    cpu = sorted(os.sched_getaffinity(0))[0]
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "cfg1.oifits")
        oifits.write_problem(path, prob)
        cmd = [O.REF_LOOP, "-d", path, "-w", str(W), "-s", repr(prob.pixellation), "-mt", "0", "-mf",
               repr(float(W * W)), "-alpha", repr(alpha), "-wavmin", "1", "-wavmax", "50000"]
        log = os.path.join(d, "ref.log")
        with open(log, "w") as fh:
            t0 = time.perf_counter()
            pr = subprocess.Popen(cmd, stdin=subprocess.DEVNULL, stdout=fh, stderr=subprocess.STDOUT, cwd=d,
                                  preexec_fn=lambda: os.sched_setaffinity(0, {cpu}))
            try:
                pr.wait(timeout=a.ref_loop_s)
            except subprocess.TimeoutExpired:
                pr.kill()
                pr.wait()
            wall = time.perf_counter() - t0
        lines = [ln for ln in open(log, errors="replace") if "Grad evals" in ln or "J = " in ln]
        tops = [ln for ln in open(log, errors="replace") if ln.startswith("Grad evals")]
        js = [ln for ln in open(log, errors="replace") if ln.startswith("J = ")]
        its = max(len(tops) - 1, 0)   # each line opens an iteration; the last one was cut off
        last = js[-1].strip() if js else ""
        gp = oifits.problem(path, W, prob.pixellation, model_type=0, model_flux=float(W * W))
    out = {"workload": f"cfg1 written as OIFITS: {W}x{W} px, {prob.npow} V2 + {prob.nbis} T3, flat prior, "
                       f"alpha {alpha}",
           "reference_cpu": {"program": "gpair1/branches/cpu_dev (oracle/_ref/gpair_cpu)", "cores": 1,
                             "iterations": its, "wall_s": wall, "iters_per_s": its / wall,
                             "last_line": last, "lines": len(lines)}}
    n_plain = max(its, 1)
    with GpairContext.from_problem(gp) as c:
        t0 = time.perf_counter()
        r, _ = c.cg_run(alpha, method=2, max_iter=n_plain, chi2r_target=0.0)
        c.L.gpair_sync(c.h)
        tw = time.perf_counter() - t0
    out["gpu_reference_rules_fp32"] = {"iterations": r.iterations, "wall_s": tw, "iters_per_s": r.iterations / tw,
                                       "chi2r": r.chi2r}
    with GpairContext.from_problem(gp) as c:
        c.set_precision(1)
        t0 = time.perf_counter()
        r, _ = c.pcg_run(alpha, precond=1, max_iter=a.max_iter, chi2r_target=1.0)
        c.L.gpair_sync(c.h)
        tw = time.perf_counter() - t0
    out["gpu_accelerated_bf16x3_to_target"] = {"iterations": r.iterations, "wall_s": tw,
                                               "iters_per_s": r.iterations / tw, "converged": bool(r.converged),
                                               "chi2r": r.chi2r}
    print(f"# reference CPU loop cfg1: {its} its in {wall:.1f} s on 1 core; GPU reference rules "
          f"{out['gpu_reference_rules_fp32']['iters_per_s']:.0f} it/s, accelerated to target "
          f"{out['gpu_accelerated_bf16x3_to_target']['wall_s']:.3f} s", file=sys.stderr, flush=True)
    return out


def fp64_errors(prob, alpha, finals):
    """relative difference between each run's reported objective and the FP64
    oracle's objective of its final image (checker only)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.Oracle()
    P = orc.unpack(prob)
    out = {}
    for name, (img, J) in finals.items():
        J64 = orc.objective(P, img.astype(np.float64), alpha)[0]
        out[name] = {"J_reported": J, "J_fp64": J64, "rel_err": abs(J - J64) / abs(J64)}
    return out


def precision_workload(a):
    """cfg3, the precision variant: 512x512 px (147,456 V2 + 344,064 T3, ring and disk
    sky).  One chi2 (forward DFT) and one data gradient (adjoint DFT) of a seeded image
    in FP32, BF16x3 and FP16 with scaling, against the FP64 oracle: max |dV| / max |V|,
    relative chi2, max |dg| / max |g|; the DFTs' mean times over 10 evaluations (HIP
    events); then cfg3_iters accelerated iterations from the prior in each precision:
    the objective the run reports (evaluated exactly at its iteration cap) and a
    separate evaluation of the final image in the working precision, both against the
    FP64 objective of that image (checker only)."""
    from gpair_amd import synth
    from gpair_amd.api import GpairContext
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    prob = synth.make_problem("cfg3", seed=a.seed)
    W, alpha = prob.W, prob.meta["alpha"]
    orc = O.Oracle()
    P = orc.unpack(prob)
    rng = np.random.default_rng(7)
    img = (prob.start.astype(np.float64) * (1.0 + rng.random(W * W))).astype(np.float32)
    t0 = time.perf_counter()
    J_o, chi2_o, S_o, vis_o, mock_o, F = orc.objective(P, img.astype(np.float64), alpha)
    gd_o = orc.gradient(P, img.astype(np.float64), alpha, vis_o, mock_o)[1]
    t_orc = time.perf_counter() - t0
    vmax, gmax = np.abs(vis_o).max(), np.abs(gd_o).max()
    out = {"workload": f"cfg3: {W}x{W} px, {prob.npow} V2 + {prob.nbis} T3, ring and disk sky, seed {a.seed}",
           "oracle_s": t_orc, "oracle_threads": orc.threads()}
    for name, p in (("fp32", 0), ("bf16x3", 1), ("fp16s", 2)):
        ctx = GpairContext.from_problem(prob, image=img)
        ctx.set_precision(p)
        chi2 = ctx.gpu_get_chi2_curr()
        vis = ctx.vis().astype(np.complex128)
        ctx.gpu_compute_data_gradient_curr()
        gd = ctx.get("dg").astype(np.float64)
        L = ctx.L
        L.gpair_timing_enable(ctx.h, 1)
        for _ in range(10):
            ctx.gpu_get_chi2_curr()
            ctx.gpu_compute_data_gradient_curr()
        tk = timing(L, ctx.h)
        L.gpair_timing_enable(ctx.h, 0)
        ctx.close()
        c2 = GpairContext.from_problem(prob)
        c2.set_precision(p)
        t1 = time.perf_counter()
        r2, _ = c2.pcg_run(alpha, precond=1, max_iter=a.cfg3_iters, chi2r_target=0.0)
        c2.L.gpair_sync(c2.h)
        wall = time.perf_counter() - t1
        fin = c2.get("ci").astype(np.float64)
        J_ex = c2.gpu_get_chi2_curr() - alpha * c2.gpu_get_entropy_curr()   # exact, working precision
        c2.close()
        J64 = orc.objective(P, fin, alpha)[0]
        fn, fms = tk["forward"]
        an, ams = tk["adjoint"]
        out[name] = {"vis_err": float(np.abs(vis - vis_o).max() / vmax), "chi2_rel_err": abs(chi2 - chi2_o) / chi2_o,
                     "grad_err": float(np.abs(gd - gd_o).max() / gmax),
                     "forward_us": fms / max(fn, 1) * 1e3, "adjoint_us": ams / max(an, 1) * 1e3,
                     "run": {"iterations": r2.iterations, "wall_s": wall, "iters_per_s": r2.iterations / wall,
                             "chi2r": r2.chi2r, "J_reported": r2.J, "J_exact": J_ex, "J_fp64": J64,
                             "J_reported_rel_err": abs(r2.J - J64) / abs(J64),
                             "J_exact_rel_err": abs(J_ex - J64) / abs(J64)}}
        print(f"# cfg3 {name}: vis {out[name]['vis_err']:.2e} chi2 {out[name]['chi2_rel_err']:.2e} "
              f"grad {out[name]['grad_err']:.2e} fwd {out[name]['forward_us']:.1f} us adj "
              f"{out[name]['adjoint_us']:.1f} us; {r2.iterations} its J err reported "
              f"{out[name]['run']['J_reported_rel_err']:.2e} exact {out[name]['run']['J_exact_rel_err']:.2e}",
              file=sys.stderr, flush=True)
    return out


def batch_workload(a):
    """cfg4: independent problems of uneven sizes (widths 64-256, 500-40,000 V2, T3 about
    2.3 V2), each run for a fixed number of accelerated-loop iterations (BF16x3): one
    gpair_pcg_run after the other, then all of them in one gpair_batch_pcg_run (one launch
    sequence per iteration for all problems, replayed from a hipGraph), and in
    gpair_batch_pcg_run_streams (one stream per problem).  Problem set-up is untimed; the
    batched runs must equal the sequential ones bit for bit."""
    from gpair_amd import synth
    from gpair_amd.api import GpairContext, batch_pcg_run
    t0 = time.perf_counter()
    probs = synth.batch_problems(a.batch)
    t_gen = time.perf_counter() - t0
    ctxs = []
    for p in probs:
        c = GpairContext.from_problem(p)
        c.set_precision(1)
        if not a.no_autotune:
            c.autotune(reps=2)
        ctxs.append(c)
    alphas = [p.meta["alpha"] for p in probs]
    starts = [c.get("ci") for c in ctxs]
    K = a.batch_iters
    key = lambda r: (r.iterations, r.J, r.chi2, r.entropy, r.forward_dfts, r.adjoint_dfts, r.ls_batches)  # noqa: E731
    t0 = time.perf_counter()
    seq = []
    for c, al in zip(ctxs, alphas):
        r, _ = c.pcg_run(al, precond=1, max_iter=K, chi2r_target=0.0)
        seq.append(key(r))
    for c in ctxs:
        c.L.gpair_sync(c.h)
    t_seq = time.perf_counter() - t0
    its = sum(s[0] for s in seq)
    out = {"problems": len(probs), "iterations_each": K, "widths": sorted({p.W for p in probs}),
           "v2_range": [min(p.npow for p in probs), max(p.npow for p in probs)],
           "t3_range": [min(p.nbis for p in probs), max(p.nbis for p in probs)],
           "sequential_s": t_seq, "iters_per_s_sequential": its / t_seq, "setup_generation_s": t_gen}
    modes = [m for m in a.batch_modes.split(",") if m in ("streams", "grouped")]
    for mode in modes:
        for c, s0 in zip(ctxs, starts):
            c.gpu_copy_image(s0)
            c.set_graphs(True)
        for c in ctxs:
            c.L.gpair_sync(c.h)
        t0 = time.perf_counter()
        res = batch_pcg_run(ctxs, alphas, max_iter=K, chi2r_target=0.0, streams=mode == "streams")
        tb = time.perf_counter() - t0
        out[mode] = {"batch_s": tb, "iters_per_s": sum(r.iterations for r in res) / tb, "speedup": t_seq / tb,
                     "identical_to_sequential": int(sum(key(r) == s for r, s in zip(res, seq)))}
        for c in ctxs:
            c.set_graphs(False)
    for c in ctxs:
        c.close()
    main_mode = "grouped" if "grouped" in out else modes[-1] if modes else None
    if main_mode:
        out["batch_s"], out["speedup"] = out[main_mode]["batch_s"], out[main_mode]["speedup"]
    print(f"# batch cfg4: {len(probs)} problems x {K} its: sequential {t_seq:.2f} s, "
          + ", ".join(f"{m} {out[m]['batch_s']:.2f} s ({out[m]['identical_to_sequential']}/{len(probs)} "
                      "bit-identical)" for m in modes), file=sys.stderr, flush=True)
    return out


def learned_workload(a):
    """cfg5: the learned start (U-Net on the dirty image and the prior, BF16 weights)
    against the plain start from the prior, on held-out problems of the training family
    (256x256, 6-telescope array): per problem the plain run to chi2/ndof = 1 (the judge)
    and the learned path (inference + the run from the learned start to the judge's
    objective; a rejected start also pays for the plain run), BF16x3."""
    from gpair_amd import learned
    if not os.path.exists(learned.WEIGHTS):
        return None
    net = learned.UNet.load(learned.WEIGHTS)
    probs = (p for _, p in learned.problems_ahead(range(learned.TEST_SEED0, learned.TEST_SEED0 + a.learned_problems)))
    t0 = time.perf_counter()
    rows = learned.evaluate(net, probs, log=lambda m: None)
    summ = learned.summarize(rows)
    summ.update({"params": net.nparam, "blend": net.blend, "seeds": [learned.TEST_SEED0,
                                                                      learned.TEST_SEED0 + a.learned_problems - 1],
                 "eval_s": time.perf_counter() - t0})
    net.close()
    print(f"# learned cfg5: {summ['plain_converged']} of {summ['problems']} held-out problems converge from the "
          f"prior; median plain {summ['median_plain_s']:.3f} s ({summ['median_plain_its']:.0f} its), learned path "
          f"{summ['median_learned_s']:.3f} s ({summ['median_path_its']:.0f} its), accepted {summ['accepted']}",
          file=sys.stderr, flush=True)
    return summ


def pixel_workload(a, prob, prec, cpu=True):
    """Single-pixel moves on the headline problem: moves scored per second by
    gpair_pixel_propose (a.moves moves per call, FP64 chi2 and entropy changes
    against the pixel state), against scoring each move by a full forward DFT and
    chi2 of the changed image on the GPU, and the FP64 port of the reference's
    incremental update (update_vis_fluxchange + vis2data + data2chi2) on the host."""
    from gpair_amd.api import PIXEL_MOVE, GpairContext
    W, nuv, npow, nbis = prob.W, prob.nuv, prob.npow, prob.nbis
    K = a.moves
    rng = np.random.default_rng(17)
    ctx = GpairContext.from_problem(prob)
    ctx.set_precision(prec)
    L = ctx.L
    img = ctx.get("ci")
    m = np.zeros(K, PIXEL_MOVE)
    m["pix_a"] = rng.integers(0, W * W, K)
    m["pix_b"] = -1
    m["flux_a"] = img[m["pix_a"]] * np.exp(rng.normal(0.0, 1.0, K)).astype(np.float32)
    two = np.arange(K) % 4 == 0   # every fourth move: two pixels
    m["pix_b"][two] = (m["pix_a"][two] + 1 + rng.integers(0, W * W - 1, two.sum())) % (W * W)
    m["flux_b"][two] = img[m["pix_b"][two]] * 1.5
    t0 = time.perf_counter()
    ctx.pixel_begin()
    L.gpair_sync(ctx.h)
    t_begin = time.perf_counter() - t0
    ctx.pixel_propose(m)   # warm-up: buffers sized for K moves
    L.gpair_timing_enable(ctx.h, 1)
    t0 = time.perf_counter()
    for _ in range(a.move_calls):
        d, s = ctx.pixel_propose(m)
    dt = time.perf_counter() - t0
    tk = {}
    for k, name in ((2, "k_px_delta"), (3, "k_px_dchi2")):
        cnt, tot = C.c_int(), C.c_double()
        L.gpair_timing_read(ctx.h, k, C.byref(cnt), C.byref(tot))
        tk[name] = tot.value / max(cnt.value, 1) * 1e3   # us per launch
    L.gpair_timing_enable(ctx.h, 0)
    # the alternative on the GPU: one forward DFT + chi2 (+ read-back) per move
    R = 50
    ctx.gpu_get_chi2_curr()
    t0 = time.perf_counter()
    for _ in range(R):
        ctx.gpu_get_chi2_curr()
    t_full = (time.perf_counter() - t0) / R
    vis_f = ctx.vis().astype(np.complex128)
    ctx.close()
    kpad = (K + 7) // 8 * 8
    # k_px_delta: reads two table columns (2 x 8 B) and writes one float2 per uv point and move
    bytes_delta = 24.0 * nuv * kpad
    # k_px_dchi2: FP64 operations per move, about 11 per V2 and 36 per T3 datum
    flop_dchi2 = (11.0 * npow + 36.0 * nbis) * kpad
    out = {"workload": f"{K} moves per call (a quarter of them two-pixel) on {W}x{W} px, {npow} V2 + {nbis} T3",
           "moves_per_s": K * a.move_calls / dt, "us_per_call": dt / a.move_calls * 1e6,
           "state_build_ms": t_begin * 1e3,
           "full_dft_moves_per_s": 1.0 / t_full, "speedup_vs_full_dft": K * a.move_calls / dt * t_full,
           "k_px_delta": {"avg_us": tk["k_px_delta"], "bytes_per_launch": bytes_delta, "bound": "hbm",
                          "achieved_GBps": bytes_delta / (tk["k_px_delta"] * 1e-6) / 1e9,
                          "frac": bytes_delta / (tk["k_px_delta"] * 1e-6) / HBM_PEAK},
           "k_px_dchi2": {"avg_us": tk["k_px_dchi2"], "fp64_flop_per_launch": flop_dchi2, "bound": "fp64 valu",
                          "achieved_tflops": flop_dchi2 / (tk["k_px_dchi2"] * 1e-6) / 1e12,
                          "frac": flop_dchi2 / (tk["k_px_dchi2"] * 1e-6) / 78.6e12}}
    print(f"# pixel moves: {out['moves_per_s']:.3g} moves/s ({out['us_per_call']:.0f} us per {K}), "
          f"k_px_delta {tk['k_px_delta']:.1f} us, k_px_dchi2 {tk['k_px_dchi2']:.1f} us; full DFT "
          f"{1.0 / t_full:.0f} moves/s", file=sys.stderr, flush=True)
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        orc = O.Oracle()
        P = orc.unpack(prob)
        ks = min(K, 2048)
        img64 = img.astype(np.float64)
        t0 = time.perf_counter()
        orc.px_propose(P, img64, vis_f, img64.sum(), m["pix_a"][:ks], m["pix_b"][:ks],
                       m["flux_a"][:ks].astype(np.float64), m["flux_b"][:ks].astype(np.float64))
        tc = time.perf_counter() - t0
        out["cpu"] = {"value": ks / tc, "unit": "moves/s", "cores": orc.threads(), "kind": "port",
                      "sample": f"{ks} moves of the same set, FP64 oracle (incremental update from the table "
                                f"identity, vis2data, data2chi2 per move), {orc.threads()} OpenMP threads"}
    return out


def timing(L, h):
    out = {}
    for k, name in ((0, "forward"), (1, "adjoint")):
        cnt, tot = C.c_int(), C.c_double()
        L.gpair_timing_read(h, k, C.byref(cnt), C.byref(tot))
        out[name] = (cnt.value, tot.value)
    return out


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    if a.calib_copy:
        calib_copy()
        return
    from gpair_amd import synth
    from gpair_amd.api import GpairContext

    prob = synth.make_problem(a.cfg, seed=a.seed)
    alpha = a.alpha if a.alpha is not None else prob.meta["alpha"]
    prec = {"fp32": 0, "bf16x3": 1, "fp16s": 2}[a.precision]
    W, nuv = prob.W, prob.nuv
    print(f"# problem {a.cfg}: W={W} nuv={nuv} nbis={prob.nbis} ndof={prob.ndof} alpha={alpha} "
          f"prior={prob.meta['prior']}", file=sys.stderr, flush=True)

    # --- headline: K timed iterations of the accelerated loop after W warm-up ones ---
    ctx = GpairContext.from_problem(prob)
    ctx.set_precision(prec)
    L = ctx.L
    tuned = None if a.no_autotune else ctx.autotune()

    def tune(c):
        if not a.no_autotune:
            c.autotune()
    if a.warmup > 0:
        ctx.pcg_run(alpha, precond=1, max_iter=a.warmup, chi2r_target=0.0)
    L.gpair_sync(ctx.h)
    # HIP events around every 8th DFT launch of the timed region (every 2nd in a run of
    # fewer than 160 steps, for at least about ten samples): events around every launch
    # cost the loop about 7 % (4,250 against 4,568 iterations/s at 256²)
    every = TIMING_EVERY if a.steps >= 160 else 2
    L.gpair_timing_enable(ctx.h, every)
    t0 = time.perf_counter()
    res, tr = ctx.pcg_run(alpha, precond=1, max_iter=a.steps, chi2r_target=0.0)
    L.gpair_sync(ctx.h)
    dt = time.perf_counter() - t0
    tk = timing(L, ctx.h)
    L.gpair_timing_enable(ctx.h, 0)
    ctx.close()
    steps_done = res.iterations
    print(f"# accelerated ({a.precision}): {steps_done} iterations in {dt:.3f} s, chi2r {res.chi2r:.5f}",
          file=sys.stderr, flush=True)

    flop = 4.0 * nuv * W * W + 8.0 * nuv * W
    fwd_n, fwd_ms = tk["forward"]
    adj_n, adj_ms = tk["adjoint"]
    avg_fwd = fwd_ms / max(fwd_n, 1) * 1e-3
    achieved = flop / avg_fwd / 1e12
    nbk = (tuned[3] if tuned and tuned[3] else min(W, 256) // 32) if prec else 0
    kname = {0: f"k_vis_fwd_f32<{W // 64}>", 1: f"k_vis_fwd_b3f<{nbk}>", 2: f"k_vis_fwd_h1f<{nbk}>"}[prec]
    # compulsory bytes of one forward: the tables (FP32: X and Y, 8 B per (uv, x) entry
    # each; BF16x3: the factored X and Y tables, 8 B per factor, W/16 + 16 factors per
    # uv point each), the image (4 B per pixel: float, or bf16 hi + lo), V out
    table_bytes = 16.0 * nuv * W if prec == 0 else 16.0 * nuv * (W // 16 + 16)
    nbytes = table_bytes + (2.0 if prec == 2 else 4.0) * W * W + 8.0 * nuv
    intensity = flop / nbytes
    traffic = read_traffic(f"hbm_traffic_W{W}_{a.precision}_pmc.txt", kname)
    if intensity < PEAK[a.precision] * 1e12 / HBM_PEAK:
        # below the machine balance: the HBM roofline binds (BF16x3)
        roofline = {"bound": "hbm", "achieved": nbytes / avg_fwd / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                    "frac": nbytes / avg_fwd / HBM_PEAK}
    else:
        roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK[a.precision], "unit": "TFLOP/s",
                    "frac": achieved / PEAK[a.precision]}
    roofline.update({"traffic": traffic, "kernel": f"{kname} (forward DFT)", "avg_us": avg_fwd * 1e6,
                     "launches": fwd_n, "timed_every": every, "flop_per_launch": flop,
                     "bytes_per_launch": nbytes, "flop_per_byte": intensity,
                     "mfma_pipe_tflops": achieved * PRODUCTS[a.precision],
                     "mfma_pipe_frac": achieved * PRODUCTS[a.precision] / PEAK[a.precision]})
    aname = {0: "k_adj_f32", 1: "k_adj_b3wf<256, 256>", 2: "k_adj_h1wf<256, 256>"}[prec] if W % 256 == 0 else (
        "k_adj_b3" if prec == 1 else "k_adj_f32")
    adj_traffic = read_traffic(f"hbm_traffic_W{W}_{a.precision}_pmc.txt", aname)
    fin_traffic = read_traffic(f"hbm_traffic_W{W}_{a.precision}_pmc.txt", "k_grad_finish_crit")
    out = {
        "metric": METRIC,
        "value": steps_done / dt,
        "unit": "CG iter/s",
        "n_gpus": 1,
        "steps": steps_done,
        "warmup": a.warmup,
        "ms_per_step": dt / steps_done * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f32", 1: "bf16x3 (f32 accumulate)", 2: "fp16 scaled (f32 accumulate)"}[prec],
        "data": "synthetic: seeded 9-telescope array, sky rendered on the grid, noise per BASELINE configs",
        "config": {"workload": f"{a.cfg}: MEM-CG reconstruction {W}x{W} px, {prob.npow} V2 + {prob.nbis} T3, "
                               f"accelerated loop (preconditioned PR-CG, batched line search), {a.precision} DFT",
                   "W": W, "nuv": nuv, "npow": prob.npow, "nbis": prob.nbis, "alpha": alpha,
                   "prior": list(prob.meta["prior"]), "seed": a.seed, "precision": a.precision},
        "roofline": roofline,
        "adjoint": {"kernel": aname, "avg_us": adj_ms / max(adj_n, 1) * 1e3, "launches": adj_n,
                    "timed_every": every,
                    "achieved_tflops": flop / (adj_ms / max(adj_n, 1) * 1e-3) / 1e12 if adj_n else None,
                    # HBM bytes per launch from the committed rocprofv3 --pmc summary: the adjoint
                    # (its split-K partials included) and the kernel that adds the partials up
                    "traffic": adj_traffic, "finish_traffic": fin_traffic,
                    "bytes_per_launch": table_bytes + 8.0 * nuv + 4.0 * W * W},
        "autotune": None if tuned is None else dict(zip(
            ("ls_chi2_blocks", "ls_entropy_blocks", "coef_blocks", "fwd_bf16x3_rows_per_group_32"), tuned)),
        "dfts_per_iter": {"forward": (res.forward_dfts) / max(steps_done, 1),
                          "adjoint": res.adjoint_dfts / max(steps_done, 1)},
    }

    # --- wall time to chi2_r = 1 ------------------------------------------------
    plain_counts = None
    finals = {}
    if not a.no_target:
        tt = {}
        for name, p in (("accelerated_fp32", 0), ("accelerated_bf16x3", 1), ("accelerated_fp16s", 2)):
            c2 = GpairContext.from_problem(prob)
            c2.set_precision(p)
            tune(c2)
            t0 = time.perf_counter()
            r2, _ = c2.pcg_run(alpha, precond=1, max_iter=a.max_iter, chi2r_target=1.0)
            c2.L.gpair_sync(c2.h)
            wall = time.perf_counter() - t0
            tt[name] = {"wall_s": wall, "iterations": r2.iterations, "iters_per_s": r2.iterations / wall,
                        "converged": bool(r2.converged), "chi2r": r2.chi2r, "J": r2.J,
                        "forward_dfts": r2.forward_dfts, "adjoint_dfts": r2.adjoint_dfts}
            finals[name] = (c2.get("ci"), r2.J)
            print(f"# {name}: {r2.iterations} its {wall:.2f} s chi2r {r2.chi2r:.6f}", file=sys.stderr, flush=True)
            c2.close()
        # spread over seeds: the time to chi2_r = 1 depends on the trajectory, which
        # rounding-level differences change
        seeds = [int(x) for x in a.spread_seeds.split(",") if x.strip()] if a.spread_seeds else []
        if seeds:
            walls, iters = [tt["accelerated_bf16x3"]["wall_s"]], [tt["accelerated_bf16x3"]["iterations"]]
            for sd in seeds:
                ps = synth.make_problem(a.cfg, seed=sd)
                c2 = GpairContext.from_problem(ps)
                c2.set_precision(1)
                tune(c2)
                t0 = time.perf_counter()
                r2, _ = c2.pcg_run(ps.meta["alpha"], precond=1, max_iter=a.max_iter, chi2r_target=1.0)
                c2.L.gpair_sync(c2.h)
                walls.append(time.perf_counter() - t0)
                iters.append(r2.iterations if r2.converged else None)
                c2.close()
            tt["accelerated_bf16x3_seeds"] = {"seeds": [a.seed] + seeds, "wall_s": walls, "iterations": iters,
                                              "median_wall_s": float(np.median(walls))}
            print(f"# accelerated_bf16x3 over seeds {[a.seed] + seeds}: {['%.2f' % w for w in walls]} s",
                  file=sys.stderr, flush=True)
        # plain loop with the reference's rules, bounded by the budget (chunks of 500 iterations)
        c3 = GpairContext.from_problem(prob)
        t0 = time.perf_counter()
        its = jev = gev = 0
        conv = False
        r3 = None
        while time.perf_counter() - t0 < a.plain_budget_s and its < a.max_iter:
            r3, _ = c3.cg_run(alpha, method=2, max_iter=500, chi2r_target=1.0)
            its += r3.iterations
            jev += r3.criterion_evals
            gev += r3.grad_evals
            if r3.converged:
                conv = True
                break
        c3.L.gpair_sync(c3.h)
        wall = time.perf_counter() - t0
        c3.close()
        tt["plain_fp32"] = {"wall_s": wall, "iterations": its, "iters_per_s": its / wall, "converged": conv,
                            "chi2r": r3.chi2r, "J": r3.J, "budget_s": a.plain_budget_s,
                            "J_evals_per_iter": jev / its, "grad_evals_per_iter": gev / its}
        plain_counts = (jev / its, gev / its)
        print(f"# plain_fp32: {its} its {wall:.2f} s chi2r {r3.chi2r:.6f} converged {conv}", file=sys.stderr,
              flush=True)
        best = min(v["wall_s"] for k, v in tt.items() if k != "plain_fp32" and v.get("converged"))
        tt["speedup_vs_plain_fp32"] = ((">=" if not conv else "") +
                                       f"{tt['plain_fp32']['wall_s'] / best:.1f}")
        out["to_target"] = tt
    if a.moves > 0 and a.move_calls > 0:
        out["pixel_moves"] = pixel_workload(a, prob, prec, cpu=not a.no_cpu)
    if a.batch > 0:
        out["batch_cfg4"] = batch_workload(a)
    if not a.no_learned and a.learned_problems > 0:
        out["learned_cfg5"] = learned_workload(a)
    if not a.no_cfg3:
        out["precision_cfg3"] = precision_workload(a)
    if not a.no_cpu:
        counts = plain_counts if plain_counts else (10.0, 2.0)
        out["cpu_baseline"] = cpu_baseline(prob, counts[0], counts[1], a.cpu_kind)
        if a.ref_loop_s > 0:
            out["reference_loop_cfg1"] = reference_loop_cfg1(a)
        if finals:
            out["err_vs_fp64"] = fp64_errors(prob, alpha, finals)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
