"""gpair_chi2_noalloc (the chi2 evaluation that allocates nothing, its
intermediates in the library's static device scratch) against gpu_get_chi2_curr
through the C ABI (MI355X only): two calls in a row, the second on a smaller
problem (the scratch keeps the first call's contents), in FP32 and BF16x3; the
same kernels and launch shapes give the same value bit for bit.  Problems beyond
the scratch's limits are refused with an error."""
import numpy as np
import pytest

from gpair_amd import synth
from gpair_amd.api import GpairContext, GpairError

pytestmark = pytest.mark.gpu


def _image(prob, seed):
    rng = np.random.default_rng(seed)
    return (prob.start.astype(np.float64) * (1.0 + rng.random(prob.W * prob.W))).astype(np.float32)


@pytest.mark.parametrize("prec", [0, 1], ids=["fp32", "bf16x3"])
def test_noalloc_chi2_matches_gpu_get_chi2(prec):
    big = synth.make_problem("cfg2", seed=3)
    small = synth.make_problem("cfg1", seed=4, W=128, ntel=7, nch=6, nsnap=5)
    for name, prob in (("cfg2 256^2", big), ("128^2", small)):
        print(f"noalloc {name} precision {prec}: W={prob.W} nuv={prob.nuv} nbis={prob.nbis}", flush=True)
        img = _image(prob, 11)
        with GpairContext.from_problem(prob, image=img) as ctx:
            ctx.set_precision(prec)
            ctx.put("ti", img[::-1].copy())
            a = ctx.chi2_noalloc("ci")
            b = ctx.gpu_get_chi2_curr()
            v_before = ctx.vis()
            c = ctx.chi2_noalloc("ti")          # another buffer; the context's visibilities stay
            assert np.array_equal(ctx.vis(), v_before)
            d = ctx.gpu_get_chi2_temp()
            print(f"  ci {a!r} vs {b!r}; ti {c!r} vs {d!r}", flush=True)
            assert a == b and c == d
            assert ctx.chi2_noalloc("ci") == b   # again: the scratch's tickets start at zero


def test_noalloc_refuses_problems_beyond_the_scratch():
    prob = synth.make_problem("cfg1", seed=5)
    with GpairContext.from_problem(prob) as ctx:
        ctx.set_precision(2)
        with pytest.raises(GpairError, match="FP32 or BF16x3"):
            ctx.chi2_noalloc("ci")
    # 2^18 + 1 uv points (V2 only) at 64^2: one more than the scratch holds
    nuv = (1 << 18) + 1
    rng = np.random.default_rng(1)
    u = rng.uniform(-3e8, 3e8, nuv).astype(np.float32)
    v = rng.uniform(-3e8, 3e8, nuv).astype(np.float32)
    data = rng.random(nuv).astype(np.float32)
    err = np.full(nuv, 50.0, np.float32)
    with GpairContext(0) as ctx:
        ctx.gpu_copy_data(data, err, nuv, np.zeros(0, np.complex64), nuv, np.zeros((0, 4), np.int64),
                          np.zeros((0, 4), np.int16), np.ones(64 * 64, np.float32), 64)
        ctx.gpair_set_uv(u, v, 0.1)
        ctx.gpu_build_kernels()
        ctx.gpu_copy_image(np.ones(64 * 64, np.float32))
        chi2 = ctx.gpu_get_chi2_curr()
        assert np.isfinite(chi2)
        with pytest.raises(GpairError, match="exceed the scratch"):
            ctx.chi2_noalloc("ci")
