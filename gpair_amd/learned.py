"""Learned start for the MEM-CG loop (include/gpair_nn.h).

A U-Net, trained and run by the library's own HIP kernels, maps two channels

  0  the dirty image: minus the chi2 gradient at the prior (the adjoint DFT of
     the data residual), divided by its RMS
  1  the log prior, log(M / mean M), clipped at -20

to two channels

  0  log of the start image over the prior mean, log(I0 / mean M)
  1  log of a per-pixel preconditioner factor (gpair_set_pfactor)

The targets come from reconstructions made by the MEM-CG loop itself (the
image I* at chi2/ndof <= 1 from the prior start, or within 1 % of it when the
loop's iteration cap comes first): log(I*/mean M) floored at
FLOOR_LOG, and log pf = GAMMA log(I*/mean M), so that the loop's image metric
z = I pf g becomes the power metric I^(1 + GAMMA) at the solution (GAMMA = 1/2
measured best on the headline problem; the plain image metric is GAMMA = 0).
The plain loop stays the fallback and the judge (run_with_start).

Training (python -m gpair_amd.learned train ...) generates the problems
(sky family learn_mix on the headline array, synth config cfg5), reconstructs
each from the prior with the accelerated loop, and fits the network with Adam
on the library's kernels; augmentation is the 180-degree rotation about the
prior's centre pixel, under which V2 is unchanged and closure phases change
sign, so the whole problem maps onto itself.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _lib

IN_CH, OUT_CH = 2, 2
FLOOR_LOG = -5.0          # log of the darkest target pixel relative to the prior mean (about 0.7 %)
ACCEPT_CHI2R = 1.01       # training targets: reconstructions within 1 % of the target chi2
GAMMA = 0.5               # power of the learned metric factor
PF_CLIP = 4.0             # |log pf| <= 4
# measured on held-out problems (seeds 900004-900013 of the family and the headline
# problem): the prediction blended 3:1 with the prior (geometric) is the start that
# converges most often and fastest; the factor channel slows every converging case
# (e.g. 1,699 instead of 467 iterations), so it is not applied by default
BLEND = 0.75
USE_PF = False
BUDGET = 6000             # iterations granted to a learned start before the fallback


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

    def save(self, path, **meta):
        np.savez(path, params=self.params, **{k: np.int64(v) for k, v in self.arch.items()},
                 **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, max_batch=1):
        z = np.load(path, allow_pickle=False)
        net = cls(int(z["width"]), base=int(z["base"]), levels=int(z["levels"]), in_ch=int(z["in_ch"]),
                  out_ch=int(z["out_ch"]), max_batch=max_batch)
        net.params = z["params"]
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


def targets(prob, final):
    """training target [2][W][W]: log(I*/mean M) floored, and GAMMA times it (log pf)."""
    W = prob.W
    m = float(prob.model.astype(np.float64).mean())
    t0 = np.maximum(np.log(np.maximum(final.astype(np.float64), 1e-30) / m), FLOOR_LOG)
    t1 = np.clip(GAMMA * t0, -PF_CLIP, PF_CLIP)
    return np.stack([t0, t1]).reshape(2, W, W).astype(np.float32)


def start_from_output(prob, y, blend=BLEND, use_pf=USE_PF):
    """start image and preconditioner factor (None unless use_pf) from the network
    output [2][W][W]: the start is the geometric blend prior^(1 - blend) * pred^blend."""
    M = prob.model.astype(np.float64)
    m = float(M.mean())
    logI = np.log(m) + np.clip(y[0].ravel().astype(np.float64), FLOOR_LOG - 2.0, 12.0)
    I0 = np.exp(blend * logI + (1.0 - blend) * np.log(np.maximum(M, 1e-30)))
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
    iterations, and, when a judge result (the plain run's gpair_pcg_result) is
    given, its objective is no worse than the judge's by more than rel_tol
    relative.  Otherwise the plain run is made (or reused from the judge) and
    returned."""
    t0 = time.perf_counter()
    ctx.gpu_copy_image(start)
    ctx.set_pfactor(pfactor)
    r, _ = ctx.pcg_run(alpha, precond=1, max_iter=min(max_iter, budget), chi2r_target=chi2r_target)
    ok = bool(r.converged) and r.chi2r <= chi2r_target
    reason = "converged" if ok else "learned start did not reach the target"
    if ok and judge is not None and r.J > judge.J + rel_tol * abs(judge.J):
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
def family_problem(seed):
    from . import synth
    return synth.make_problem("cfg5", seed=seed)


def reconstruct(prob, precision=1, max_iter=15000):
    """features at the prior, and the accelerated loop's reconstruction from the prior."""
    from .api import GpairContext
    alpha = prob.meta["alpha"]
    with GpairContext.from_problem(prob) as ctx:
        ctx.set_precision(precision)
        x = features(ctx, prob)
        ctx.gpu_copy_image(prob.start)
        t0 = time.perf_counter()
        r, _ = ctx.pcg_run(alpha, precond=1, max_iter=max_iter, chi2r_target=1.0)
        wall = time.perf_counter() - t0
        final = ctx.get("ci")
    return x, final, r, wall


def rot180(a):
    """180-degree rotation about pixel (W/2, W/2) of the last two axes."""
    return np.roll(a[..., ::-1, ::-1], 1, axis=(-2, -1))


def generate(seeds, precision=1, max_iter=12000, log=print):
    """problems are built on host threads (FP64 numpy DFTs) ahead of the GPU loop."""
    from concurrent.futures import ThreadPoolExecutor
    X, F, info = [], [], []
    seeds = list(seeds)
    pool = ThreadPoolExecutor(max_workers=3)
    futs = [pool.submit(family_problem, s) for s in seeds[:3]]
    for i, s in enumerate(seeds):
        prob = futs[i].result()
        if i + 3 < len(seeds):
            futs.append(pool.submit(family_problem, seeds[i + 3]))
        x, final, r, wall = reconstruct(prob, precision, max_iter)
        log(f"sample {i} seed {s}: sky {prob.meta['components'][0]['kind']} its {r.iterations} "
            f"{wall:.2f}s chi2r {r.chi2r:.5f} converged {bool(r.converged)}")
        if not (r.chi2r <= ACCEPT_CHI2R):
            continue
        X.append(x)
        F.append(final)
        info.append((s, r.iterations, wall, r.J))
    pool.shutdown()
    return np.stack(X), np.stack(F), info


def train(net, X, T, epochs=40, batch=8, lr=1e-3, seed=0, log=print):
    """Adam on random batches with the rotation augmentation; cosine learning rate."""
    rng = np.random.default_rng(seed)
    n = X.shape[0]
    W = X.shape[-1]
    xs = DeviceArray(batch * IN_CH * W * W)
    ts = DeviceArray(batch * OUT_CH * W * W)
    steps = max(1, epochs * n // batch)
    hist = []
    for k in range(steps):
        idx = rng.choice(n, size=batch, replace=n < batch)
        xb, tb = X[idx].copy(), T[idx].copy()
        flip = rng.random(batch) < 0.5
        xb[flip], tb[flip] = rot180(xb[flip]), rot180(tb[flip])
        xs.upload(xb)
        ts.upload(tb)
        lr_k = lr * 0.5 * (1.0 + np.cos(np.pi * k / steps))
        loss = net.train_step(xs.at(0), ts.at(0), batch, max(lr_k, lr * 1e-3))
        hist.append(loss)
        if k % 100 == 0 or k == steps - 1:
            log(f"step {k}/{steps} loss {np.mean(hist[-100:]):.5f} lr {lr_k:.2e}")
    xs.close()
    ts.close()
    return hist


def predict(net, ctx, prob):
    """start image and preconditioner factor for a problem held by ctx."""
    x = features(ctx, prob)
    y = net.forward(x[None])[0]
    return start_from_output(prob, y)


def evaluate(net, seeds, precision=1, max_iter=15000, log=print, problems=None, autotune=True):
    """plain start (the judge) against the learned start, per problem."""
    from .api import GpairContext
    rows = []
    probs = problems if problems is not None else [family_problem(s) for s in seeds]
    for prob in probs:
        alpha = prob.meta["alpha"]
        with GpairContext.from_problem(prob) as ctx:
            ctx.set_precision(precision)
            if autotune:
                ctx.autotune()
            t0 = time.perf_counter()
            rp, _ = ctx.pcg_run(alpha, precond=1, max_iter=max_iter, chi2r_target=1.0)
            ctx.L.gpair_sync(ctx.h)
            wp = time.perf_counter() - t0
            t0 = time.perf_counter()
            start, pf = predict(net, ctx, prob)
            t_inf = time.perf_counter() - t0
            out = run_with_start(ctx, prob, alpha, start, pf, max_iter=max_iter, judge=rp)
            ctx.set_pfactor(None)
        lr_ = out.learned_result
        cost = t_inf + out.wall_s + (0.0 if out.accepted else wp)
        row = dict(cfg=prob.meta["cfg"], seed=int(prob.meta["seed"]), plain_its=rp.iterations, plain_s=wp,
                   plain_J=rp.J, plain_conv=bool(rp.converged), learned_its=lr_.iterations, learned_s=cost,
                   learned_J=lr_.J, learned_conv=bool(lr_.converged), accepted=out.accepted, infer_s=t_inf,
                   reason=out.reason)
        log(f"{row['cfg']} seed {row['seed']}: plain {rp.iterations} its {wp:.2f} s J {rp.J:.2f} conv "
            f"{bool(rp.converged)} | learned {lr_.iterations} its, path {cost:.2f} s, J {lr_.J:.2f}, accepted "
            f"{out.accepted} ({out.reason})")
        rows.append(row)
    return rows


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_cfg5.npz")


def save_data(path, X, F, info):
    """dirty-image channel and floored log reconstruction as float16 (the prior
    channel is the same for every problem of the family and is stored once)."""
    np.savez_compressed(path, dirty=X[:, 0].astype(np.float16), logprior=X[0, 1].astype(np.float32),
                        logF=np.maximum(np.log(np.maximum(F.astype(np.float64), 1e-30)), FLOOR_LOG)
                        .astype(np.float16), info=np.array(info, np.float64))


def load_data(path):
    z = np.load(path, allow_pickle=False)
    d = z["dirty"].astype(np.float32)
    n, W = d.shape[0], d.shape[-1]
    X = np.stack([d, np.broadcast_to(z["logprior"], d.shape)], axis=1).astype(np.float32)
    lf = z["logF"].astype(np.float32).reshape(n, W, W)
    T = np.stack([lf, np.clip(GAMMA * lf, -PF_CLIP, PF_CLIP)], axis=1).astype(np.float32)
    return X, T, z["info"]


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m gpair_amd.learned")
    ap.add_argument("mode", choices=["generate", "train", "eval"])
    ap.add_argument("--n", type=int, default=240, help="training problems")
    ap.add_argument("--seed0", type=int, default=100000, help="first training seed")
    ap.add_argument("--data", default="gpurun_out/learn_data.npz")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--base", type=int, default=16)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--eval-seeds", default="900000:900008")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)

    def log(m):
        print(m, flush=True)

    if a.mode in ("generate", "train"):
        if a.mode == "generate" or not os.path.exists(a.data):
            t0 = time.perf_counter()
            X, F, info = generate(range(a.seed0, a.seed0 + a.n), log=log)
            save_data(a.data, X, F, info)
            log(f"saved {X.shape[0]} samples to {a.data} ({time.perf_counter() - t0:.0f} s)")
        if a.mode == "train":
            X, T, info = load_data(a.data)
            W = X.shape[-1]
            net = UNet(W, base=a.base, levels=a.levels, max_batch=a.batch, seed=1)
            log(f"training {net.nparam} parameters on {X.shape[0]} samples")
            t0 = time.perf_counter()
            hist = train(net, X, T, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
            log(f"trained in {time.perf_counter() - t0:.1f} s, final loss {np.mean(hist[-50:]):.5f}")
            os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
            net.save(a.out, n_train=X.shape[0], epochs=a.epochs, final_loss=float(np.mean(hist[-50:])))
            log(f"saved {a.out}")
    if a.mode == "eval":
        lo, hi = (int(v) for v in a.eval_seeds.split(":"))
        net = UNet.load(a.out)
        rows = evaluate(net, range(lo, hi), log=log)
        if a.json:
            with open(a.json, "w") as f:
                json.dump(rows, f)


if __name__ == "__main__":
    main()
