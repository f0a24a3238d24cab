"""Learned start for the MEM-CG loop (include/gpair_nn.h).

A U-Net, trained and run by the library's own HIP kernels, maps two channels

  0  the dirty image: minus the chi2 gradient at the prior (the adjoint DFT of
     the data residual), divided by its RMS
  1  the log prior, log(M / mean M), clipped at -20

to two channels

  0  the log ratio of the start image to the prior, log(I0 / M)
  1  log of a per-pixel preconditioner factor (gpair_set_pfactor)

The targets come from reconstructions made by the MEM-CG loop itself (the
image I* at chi2/ndof <= 1 from the prior start, or within 1 % of it when the
loop's iteration cap comes first): log(I*/M) clipped to +-RCLIP (zero wherever
the reconstruction keeps the prior), and log pf = GAMMA log(I*/M), so that the
loop's image metric z = I pf g becomes the power metric I^(1 + GAMMA) M^-GAMMA at
the solution (the plain image metric is GAMMA = 0).
The plain loop stays the fallback and the judge (run_with_start).

Training (python -m gpair_amd.learned all ...) generates the problems (sky
family learn_mix on the 256x256 grid sampled by the 6-telescope array, synth
config cfg5), reconstructs each from the prior with the accelerated loop, many
problems per gpair_batch_pcg_run, and fits the network with Adam on the
library's kernels (BF16 operands, FP32 accumulation); augmentation is the
180-degree rotation about the prior's centre pixel, under which V2 is unchanged
and closure phases change sign, so the whole problem maps onto itself.  The
stored parameters are BF16, as the convolutions use them.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

IN_CH, OUT_CH = 2, 2
RCLIP = 8.0               # |log(I*/M)| <= RCLIP in the targets and the start
ACCEPT_CHI2R = 1.01       # training targets: reconstructions within 1 % of the target chi2
GAMMA = 0.5               # power of the learned metric factor
PF_CLIP = 4.0             # |log pf| <= 4
# start = prior^(1 - BLEND) * prediction^BLEND (geometric); a trained network stores
# the blend chosen on validation problems, BLEND is the default for one that does not;
# the factor channel slowed every converging case of the headline array (e.g. 1,699
# instead of 467 iterations), so it is not applied by default
BLEND = 0.75
USE_PF = False
BUDGET = 500              # iterations granted to a learned start before the fallback


def _p(a):
    return C.c_void_p(a.ctypes.data)


class DeviceArray:
    """float32 device buffer from gpair_nn_alloc (freed on close)."""

    def __init__(self, n, data=None):
        self.L = _lib.load()
        self.n = int(n)
        self.ptr = self.L.gpair_nn_alloc(self.n)
        if not self.ptr:
            raise MemoryError(f"gpair_nn_alloc({self.n})")
        if data is not None:
            self.upload(data)

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a, np.float32).ravel()
        if offset + a.size > self.n:
            raise ValueError("upload beyond the buffer")
        rc = self.L.gpair_nn_upload(C.c_void_p(self.ptr + 4 * offset), _p(a), a.size)
        if rc:
            raise RuntimeError(f"gpair_nn_upload: {rc}")

    def download(self, n=None, offset=0):
        n = self.n - offset if n is None else n
        out = np.empty(n, np.float32)
        rc = self.L.gpair_nn_download(_p(out), C.c_void_p(self.ptr + 4 * offset), n)
        if rc:
            raise RuntimeError(f"gpair_nn_download: {rc}")
        return out

    def at(self, offset):
        return C.c_void_p(self.ptr + 4 * offset)

    def close(self):
        if self.ptr:
            self.L.gpair_nn_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UNet:
    """gpair_unet: create / forward / backward / train_step / save / load."""

    def __init__(self, width, base=16, levels=4, in_ch=IN_CH, out_ch=OUT_CH, max_batch=8, seed=1):
        self.L = _lib.load()
        self.arch = dict(width=int(width), base=int(base), levels=int(levels), in_ch=int(in_ch),
                         out_ch=int(out_ch))
        self.max_batch = int(max_batch)
        h = C.c_void_p()
        rc = self.L.gpair_unet_create(C.byref(h), in_ch, out_ch, base, levels, width, max_batch, seed)
        if rc:
            raise RuntimeError(f"gpair_unet_create: status {rc}")
        self.h = h
        self.nparam = int(self.L.gpair_unet_num_params(h))

    def close(self):
        if self.h:
            self.L.gpair_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: status {rc}")

    @property
    def params(self):
        out = np.empty(self.nparam, np.float32)
        self._check(self.L.gpair_unet_get_params(self.h, _p(out), self.nparam), "get_params")
        return out

    @params.setter
    def params(self, v):
        v = np.ascontiguousarray(v, np.float32)
        self._check(self.L.gpair_unet_set_params(self.h, _p(v), v.size), "set_params")

    def grads(self):
        out = np.empty(self.nparam, np.float32)
        self._check(self.L.gpair_unet_get_grads(self.h, _p(out), self.nparam), "get_grads")
        return out

    def forward(self, x):
        """x [N][in][W][W] float32 (host) -> y [N][out][W][W]."""
        x = np.ascontiguousarray(x, np.float32)
        N = x.shape[0]
        y = np.empty((N, self.arch["out_ch"], self.arch["width"], self.arch["width"]), np.float32)
        for a in range(0, N, self.max_batch):
            b = min(N, a + self.max_batch)
            xa = np.ascontiguousarray(x[a:b])
            ya = np.empty((b - a,) + y.shape[1:], np.float32)
            self._check(self.L.gpair_unet_forward(self.h, _p(xa), b - a, _p(ya)), "forward")
            y[a:b] = ya
        return y

    def backward(self, dy):
        dy = np.ascontiguousarray(dy, np.float32)
        self._check(self.L.gpair_unet_backward(self.h, _p(dy), dy.shape[0]), "backward")

    def train_step(self, xdev, tdev, N, lr):
        """one Adam step on N samples already on the device (pointers)."""
        loss = C.c_double()
        self._check(self.L.gpair_unet_train_step(self.h, xdev, tdev, int(N), C.c_float(lr), C.byref(loss)),
                    "train_step")
        return loss.value

    @classmethod
    def load(cls, path, max_batch=1):
        """a network saved by main(): BF16 parameters (params_bf16) or FP32 ones
        (params); the start blend chosen on validation problems when stored."""
        z = np.load(path, allow_pickle=False)
        net = cls(int(z["width"]), base=int(z["base"]), levels=int(z["levels"]), in_ch=int(z["in_ch"]),
                  out_ch=int(z["out_ch"]), max_batch=max_batch)
        net.params = from_bf16_bits(z["params_bf16"]) if "params_bf16" in z.files else z["params"]
        net.blend = float(z["blend"]) if "blend" in z.files else BLEND
        return net


# ---------------------------------------------------------------------------
# features and targets
# ---------------------------------------------------------------------------
def features(ctx, prob):
    """network input [2][W][W] for a problem (ctx: a GpairContext holding it)."""
    W = prob.W
    M = prob.model.astype(np.float64)
    ctx.gpu_copy_image(prob.start)
    ctx.gpu_get_chi2_curr()
    ctx.gpu_compute_data_gradient_curr()
    g = ctx.get("dg").astype(np.float64)
    d = -g / (np.sqrt(np.mean(g * g)) + 1e-30)
    lp = np.maximum(np.log(np.maximum(M, 1e-300) / M.mean()), -20.0)
    return np.stack([d, lp]).reshape(2, W, W).astype(np.float32)


def log_ratio(prob, final):
    """log(I*/M) per pixel, clipped to +-RCLIP: 0 wherever the reconstruction keeps the
    prior (the background of a MEM image), so the network learns only the departures."""
    M = np.maximum(prob.model.astype(np.float64), 1e-30)
    return np.clip(np.log(np.maximum(final.astype(np.float64), 1e-30) / M), -RCLIP, RCLIP)


def targets(prob, final):
    """training target [2][W][W]: log(I*/M) clipped, and GAMMA times it (log pf)."""
    t0 = log_ratio(prob, final)
    t1 = np.clip(GAMMA * t0, -PF_CLIP, PF_CLIP)
    return np.stack([t0, t1]).reshape(2, prob.W, prob.W).astype(np.float32)


def start_from_output(prob, y, blend=BLEND, use_pf=USE_PF):
    """start image and preconditioner factor (None unless use_pf) from the network
    output [2][W][W]: the start is M exp(blend * predicted log(I*/M)), the geometric
    blend prior^(1 - blend) * prediction^blend."""
    M = prob.model.astype(np.float64)
    I0 = M * np.exp(blend * np.clip(y[0].ravel().astype(np.float64), -RCLIP, RCLIP))
    pf = np.exp(np.clip(y[1].ravel().astype(np.float64), -PF_CLIP, PF_CLIP)).astype(np.float32) if use_pf else None
    return np.maximum(I0, 1e-8).astype(np.float32), pf


@dataclass
class StartOutcome:
    accepted: bool
    reason: str
    result: object          # gpair_pcg_result of the run that is returned
    wall_s: float
    learned_result: object = None
    plain_result: object = None


def run_with_start(ctx, prob, alpha, start, pfactor=None, max_iter=30000, chi2r_target=1.0, judge=None,
                   rel_tol=1e-4, budget=BUDGET):
    """Runs the accelerated loop from a learned start; the plain start (the prior,
    no factor) is the fallback and the judge.

    Acceptance: the learned run reaches chi2/ndof <= chi2r_target within `budget`
    iterations and, when a judge result (the plain run's gpair_pcg_result) is
    given, an objective within rel_tol of the judge's or below it: the run goes
    on past chi2/ndof = 1 until J <= (1 + rel_tol) J_judge (gpair_pcg_params
    J_target), so an accepted start costs what it takes to reach a solution as
    good as the plain run's.  Otherwise the plain run is made (or reused from the
    judge) and returned."""
    t0 = time.perf_counter()
    ctx.gpu_copy_image(start)
    ctx.set_pfactor(pfactor)
    jt = judge.J + rel_tol * abs(judge.J) if judge is not None else 0.0
    r, _ = ctx.pcg_run(alpha, precond=1, max_iter=min(max_iter, budget), chi2r_target=chi2r_target,
                       J_target=max(jt, 0.0))
    ok = bool(r.converged) and r.chi2r <= chi2r_target
    reason = "converged" if ok else "learned start did not reach the target"
    if ok and judge is not None and r.J > jt:
        ok = False
        reason = f"objective {r.J:.8e} above the plain run's {judge.J:.8e} by more than {rel_tol:g}"
    if ok:
        return StartOutcome(True, reason, r, time.perf_counter() - t0, learned_result=r, plain_result=judge)
    ctx.set_pfactor(None)
    if judge is not None:
        return StartOutcome(False, reason, judge, time.perf_counter() - t0, learned_result=r, plain_result=judge)
    ctx.gpu_copy_image(prob.start)
    rp, _ = ctx.pcg_run(alpha, precond=1, max_iter=max_iter, chi2r_target=chi2r_target)
    return StartOutcome(False, reason, rp, time.perf_counter() - t0, learned_result=r, plain_result=rp)


# ---------------------------------------------------------------------------
# data generation, training, evaluation
# ---------------------------------------------------------------------------
CFG = "cfg5"              # 256x256 grid, 6-telescope array (synth.CONFIGS)
TRAIN_SEED0 = 100000      # training problems: seeds from here
VAL_SEED0 = 800000        # validation problems (loss curve, choice of the blend)
TEST_SEED0 = 900000       # held-out problems (evaluation only)


def family_problem(seed):
    from . import synth
    return synth.make_problem(CFG, seed=seed)


def problems_ahead(seeds, workers=6):
    """(seed, problem) in seed order, built on host threads ahead of use (the FP64
    numpy DFTs release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    seeds = list(seeds)
    ahead = 4 * workers
    with ThreadPoolExecutor(max_workers=workers) as pool:
        futs = {k: pool.submit(family_problem, seeds[k]) for k in range(min(ahead, len(seeds)))}
        for i, s in enumerate(seeds):
            prob = futs.pop(i).result()
            if i + ahead < len(seeds):
                futs[i + ahead] = pool.submit(family_problem, seeds[i + ahead])
            yield s, prob


def rot180(a):
    """180-degree rotation about pixel (W/2, W/2) of the last two axes."""
    return np.roll(a[..., ::-1, ::-1], 1, axis=(-2, -1))


def to_bf16_bits(a):
    """float32 -> bfloat16 bit patterns, rounded to nearest even (finite values)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


@dataclass
class Dataset:
    dirty: np.ndarray       # [n][W][W] float16: the dirty-image channel
    logR: np.ndarray        # [n][W][W] float16: log(I*/M) clipped to +-RCLIP
    logprior: np.ndarray    # [W][W] float32: the prior channel (one prior for the family)
    info: np.ndarray        # [n][4]: seed, iterations, J, chi2/ndof of the reconstruction

    def __len__(self):
        return self.dirty.shape[0]

    def inputs(self, idx):
        d = self.dirty[idx].astype(np.float32)
        return np.stack([d, np.broadcast_to(self.logprior, d.shape)], axis=1)

    def targets(self, idx):
        t0 = self.logR[idx].astype(np.float32)
        return np.stack([t0, np.clip(GAMMA * t0, -PF_CLIP, PF_CLIP)], axis=1)


def generate(seeds, precision=1, max_iter=4000, chunk=16, log=print):
    """Training pairs from reconstructions made by the plain method (the accelerated
    loop from the prior to chi2/ndof <= 1), `chunk` problems per gpair_batch_pcg_run.
    Problems that end above ACCEPT_CHI2R (the loop's cap came first) are dropped."""
    import itertools
    from .api import GpairContext, batch_pcg_run
    dirty, logR, info = [], [], []
    logprior = None
    t0 = time.perf_counter()
    seen = 0
    src = problems_ahead(seeds)
    while True:
        group = list(itertools.islice(src, chunk))
        if not group:
            break
        ctxs = []
        try:
            feats = []
            for _, prob in group:
                ctx = GpairContext.from_problem(prob)
                ctxs.append(ctx)
                ctx.set_precision(precision)
                feats.append(features(ctx, prob))
                ctx.gpu_copy_image(prob.start)
            res = batch_pcg_run(ctxs, [p.meta["alpha"] for _, p in group], max_iter=max_iter, chi2r_target=1.0)
            for (s, prob), ctx, x, r in zip(group, ctxs, feats, res):
                if logprior is None:
                    logprior = x[1].copy()
                if not r.chi2r <= ACCEPT_CHI2R:
                    continue
                dirty.append(x[0].astype(np.float16))
                logR.append(log_ratio(prob, ctx.get("ci")).reshape(prob.W, prob.W).astype(np.float16))
                info.append((s, r.iterations, r.J, r.chi2r))
        finally:
            for c in ctxs:
                c.close()
        seen += len(group)
        if seen % (16 * chunk) == 0 or seen == len(seeds):
            log(f"generated {seen}/{len(seeds)} problems, kept {len(dirty)}, {time.perf_counter() - t0:.0f} s")
    return Dataset(np.stack(dirty), np.stack(logR), logprior, np.array(info, np.float64))


def batch_mse(net, data, idx, batch=16):
    """mean squared error of the network on data[idx] (forward only)."""
    tot, cnt = 0.0, 0
    for a in range(0, len(idx), batch):
        j = idx[a:a + batch]
        y = net.forward(data.inputs(j)).astype(np.float64)
        t = data.targets(j).astype(np.float64)
        tot += float(((y - t) ** 2).sum())
        cnt += y.size
    return tot / max(cnt, 1)


def train(net, data, val=None, epochs=20, batch=16, lr=1e-3, seed=0, log=print):
    """Adam on shuffled batches with the rotation augmentation and a cosine learning
    rate; returns the learning curve rows (epoch, steps, mean training loss of the
    epoch, validation loss)."""
    rng = np.random.default_rng(seed)
    n = len(data)
    W = data.dirty.shape[-1]
    xs = DeviceArray(batch * IN_CH * W * W)
    ts = DeviceArray(batch * OUT_CH * W * W)
    per = max(1, n // batch)
    total = epochs * per
    curve = []
    k = 0
    vidx = np.arange(len(val)) if val is not None else None
    t0 = time.perf_counter()
    try:
        for ep in range(epochs):
            perm = rng.permutation(n)
            losses = []
            for b in range(per):
                idx = np.sort(perm[b * batch:(b + 1) * batch])
                xb, tb = data.inputs(idx), data.targets(idx)
                flip = rng.random(len(idx)) < 0.5
                xb[flip], tb[flip] = rot180(xb[flip]), rot180(tb[flip])
                xs.upload(xb)
                ts.upload(tb)
                lr_k = lr * 0.5 * (1.0 + np.cos(np.pi * k / total))
                losses.append(net.train_step(xs.at(0), ts.at(0), len(idx), max(lr_k, lr * 1e-3)))
                k += 1
            vl = batch_mse(net, val, vidx, batch) if val is not None else float("nan")
            curve.append((ep + 1, k, float(np.mean(losses)), vl))
            log(f"epoch {ep + 1}/{epochs}: {k} steps, training loss {np.mean(losses):.5f}, validation loss "
                f"{vl:.5f} ({time.perf_counter() - t0:.0f} s)")
    finally:
        xs.close()
        ts.close()
    return curve


def predict(net, ctx, prob, blend=None):
    """start image and preconditioner factor for a problem held by ctx."""
    x = features(ctx, prob)
    y = net.forward(x[None])[0]
    return start_from_output(prob, y, blend=getattr(net, "blend", BLEND) if blend is None else blend)


def evaluate(net, problems, precision=1, max_iter=4000, blend=None, log=print):
    """Per problem: the plain start (the judge) and the learned start, each timed to
    chi2/ndof <= 1 after an untimed one-iteration run that sets up the context (split
    tables, graphs).  The learned path's cost is inference plus its run, plus the plain
    run when the start is rejected."""
    from .api import GpairContext
    rows = []
    for prob in problems:
        alpha = prob.meta["alpha"]
        with GpairContext.from_problem(prob) as ctx:
            ctx.set_precision(precision)
            ctx.pcg_run(alpha, precond=1, max_iter=1, chi2r_target=0.0)
            ctx.gpu_copy_image(prob.start)
            t0 = time.perf_counter()
            rp, _ = ctx.pcg_run(alpha, precond=1, max_iter=max_iter, chi2r_target=1.0)
            ctx.L.gpair_sync(ctx.h)
            wp = time.perf_counter() - t0
            t0 = time.perf_counter()
            start, pf = predict(net, ctx, prob, blend)
            t_inf = time.perf_counter() - t0
            out = run_with_start(ctx, prob, alpha, start, pf, max_iter=max_iter, judge=rp)
            ctx.set_pfactor(None)
        lr_ = out.learned_result
        cost = t_inf + out.wall_s + (0.0 if out.accepted else wp)
        rows.append(dict(cfg=prob.meta["cfg"], seed=int(prob.meta["seed"]), plain_its=rp.iterations, plain_s=wp,
                         plain_J=rp.J, plain_conv=bool(rp.converged), learned_its=lr_.iterations,
                         path_its=lr_.iterations + (0 if out.accepted else rp.iterations), learned_s=cost,
                         learned_J=lr_.J, learned_conv=bool(lr_.converged), accepted=out.accepted, infer_s=t_inf,
                         reason=out.reason))
        r = rows[-1]
        if len(rows) <= 8 or len(rows) % 64 == 0:
            log(f"{r['cfg']} seed {r['seed']}: plain {r['plain_its']} its {wp:.3f} s | learned {r['learned_its']} "
                f"its, path {cost:.3f} s, accepted {out.accepted} ({out.reason})")
    return rows


def summarize(rows):
    """medians and totals of an evaluation; objective differences of the converged
    learned runs relative to the judge's."""
    conv = [r for r in rows if r["plain_conv"]]
    dJ = np.array([(r["learned_J"] - r["plain_J"]) / abs(r["plain_J"]) for r in conv if r["learned_conv"]])
    med = lambda k: float(np.median([r[k] for r in conv])) if conv else float("nan")  # noqa: E731
    return dict(problems=len(rows), plain_converged=len(conv), accepted=int(sum(r["accepted"] for r in conv)),
                median_plain_s=med("plain_s"), median_learned_s=med("learned_s"),
                median_plain_its=med("plain_its"), median_path_its=med("path_its"),
                total_plain_s=float(sum(r["plain_s"] for r in conv)),
                total_learned_s=float(sum(r["learned_s"] for r in conv)),
                dJ_rel_quantiles=[float(q) for q in np.quantile(dJ, [0.0, 0.1, 0.5, 0.9, 1.0])] if dJ.size else [],
                dJ_within_1e4=int(np.sum(np.abs(dJ) <= 1e-4)), dJ_below=int(np.sum(dJ < 0.0)))


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_cfg5.npz")


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m gpair_amd.learned")
    ap.add_argument("mode", choices=["all", "more", "eval"],
                    help="all: generate, train from scratch, evaluate; more: go on training the saved "
                         "network on --n new problems from --train-seed0; eval: evaluate the saved network")
    ap.add_argument("--train-seed0", type=int, default=TRAIN_SEED0, help="first training seed (mode more)")
    ap.add_argument("--n", type=int, default=4000, help="training problems")
    ap.add_argument("--n-val", type=int, default=128, help="validation problems")
    ap.add_argument("--n-test", type=int, default=512, help="held-out problems")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--base", type=int, default=16)
    ap.add_argument("--levels", type=int, default=5)
    ap.add_argument("--blends", default="0.5,0.75,1.0", help="start blends tried on validation problems")
    ap.add_argument("--n-blend", type=int, default=32, help="validation problems per blend trial")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    report = {}
    if a.mode == "more":
        # further training on fresh problems of the same family: the generated data
        # cannot outlive one job, so a larger training set is seen in rounds; the
        # record keeps every round's seeds and kept reconstructions
        prev = np.load(a.out, allow_pickle=False)
        t0 = time.perf_counter()
        data = generate(range(a.train_seed0, a.train_seed0 + a.n), log=log)
        val = generate(range(VAL_SEED0, VAL_SEED0 + a.n_val), log=log)
        t_gen = time.perf_counter() - t0
        log(f"data: {len(data)} training and {len(val)} validation reconstructions in {t_gen:.0f} s")
        net = UNet.load(a.out, max_batch=a.batch)
        t0 = time.perf_counter()
        curve = train(net, data, val, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.train_seed0, log=log)
        t_train = time.perf_counter() - t0
        net.params = from_bf16_bits(to_bf16_bits(net.params))
        trials = {}
        for b in (float(v) for v in a.blends.split(",")):
            probs = [family_problem(s) for s in range(VAL_SEED0, VAL_SEED0 + a.n_blend)]
            trials[b] = summarize(evaluate(net, probs, blend=b, log=lambda m: None))
            log(f"blend {b}: {trials[b]}")
        blend = min(trials, key=lambda b: (trials[b]["median_learned_s"], -trials[b]["accepted"]))
        rnd = int(prev["rounds"]) + 1 if "rounds" in prev.files else 2
        info = np.concatenate([prev["train_info"], data.info])
        # learning curve rows: round, epoch, steps, training loss, validation loss
        old_curve = prev["curve"] if prev["curve"].shape[1] == 5 else np.concatenate(
            [np.ones((len(prev["curve"]), 1)), prev["curve"]], axis=1)
        new_curve = np.concatenate([np.full((len(curve), 1), float(rnd)), np.array(curve, np.float64)], axis=1)
        np.savez(a.out, params_bf16=to_bf16_bits(net.params), **{k: np.int64(v) for k, v in net.arch.items()},
                 blend=np.float64(blend), n_train=np.int64(len(info)), n_val=np.int64(len(val)),
                 epochs=np.int64(int(prev["epochs"]) + a.epochs), batch=np.int64(a.batch), lr=np.float64(a.lr),
                 curve=np.concatenate([old_curve, new_curve]), train_info=info,
                 gen_s=np.float64(float(prev["gen_s"]) + t_gen), train_s=np.float64(float(prev["train_s"]) + t_train),
                 rounds=np.int64(rnd))
        log(f"saved {a.out} (round {rnd}: {len(info)} reconstructions in all, blend {blend})")
        report.update(n_train=len(info), round=rnd, gen_s=t_gen, train_s=t_train, curve=curve,
                      blend_trials={str(k): v for k, v in trials.items()}, blend=blend, params=net.nparam)
        net.close()
    if a.mode == "all":
        t0 = time.perf_counter()
        data = generate(range(TRAIN_SEED0, TRAIN_SEED0 + a.n), log=log)
        val = generate(range(VAL_SEED0, VAL_SEED0 + a.n_val), log=log)
        t_gen = time.perf_counter() - t0
        log(f"data: {len(data)} training and {len(val)} validation reconstructions in {t_gen:.0f} s")
        W = data.dirty.shape[-1]
        net = UNet(W, base=a.base, levels=a.levels, max_batch=a.batch, seed=1)
        log(f"training {net.nparam} parameters")
        t0 = time.perf_counter()
        curve = train(net, data, val, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
        t_train = time.perf_counter() - t0
        # the stored network is the one the kernels run: BF16 parameters
        net.params = from_bf16_bits(to_bf16_bits(net.params))
        trials = {}
        for b in (float(v) for v in a.blends.split(",")):
            probs = [family_problem(s) for s in range(VAL_SEED0, VAL_SEED0 + a.n_blend)]
            trials[b] = summarize(evaluate(net, probs, blend=b, log=lambda m: None))
            log(f"blend {b}: {trials[b]}")
        blend = min(trials, key=lambda b: (trials[b]["median_learned_s"], -trials[b]["accepted"]))
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.savez(a.out, params_bf16=to_bf16_bits(net.params), **{k: np.int64(v) for k, v in net.arch.items()},
                 blend=np.float64(blend), n_train=np.int64(len(data)), n_val=np.int64(len(val)),
                 epochs=np.int64(a.epochs), batch=np.int64(a.batch), lr=np.float64(a.lr),
                 curve=np.array(curve, np.float64), train_info=data.info, gen_s=np.float64(t_gen),
                 train_s=np.float64(t_train))
        log(f"saved {a.out} (blend {blend}, {t_train:.0f} s of training)")
        report.update(n_train=len(data), n_val=len(val), gen_s=t_gen, train_s=t_train, curve=curve,
                      blend_trials={str(k): v for k, v in trials.items()}, blend=blend, params=net.nparam)
        net.close()
    net = UNet.load(a.out, max_batch=1)
    probs = problems_ahead(range(TEST_SEED0, TEST_SEED0 + a.n_test))
    rows = evaluate(net, (p for _, p in probs), log=log)
    report["test"] = summarize(rows)
    log(f"held-out: {report['test']}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump(dict(report, rows=rows), f)


if __name__ == "__main__":
    main()
