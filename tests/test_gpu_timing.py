"""DFT timing with HIP events (MI355X only): events around every launch, or around
every N-th launch of each kernel, over identical runs of the accelerated loop."""
import ctypes as C
import math

import pytest

from gpair_amd import synth
from gpair_amd.api import GpairContext

pytestmark = pytest.mark.gpu


def _timed_run(prob, every):
    with GpairContext.from_problem(prob) as ctx:
        ctx.set_precision(1)
        assert ctx.L.gpair_timing_enable(ctx.h, every) == 0
        res, _ = ctx.pcg_run(prob.meta["alpha"], precond=1, max_iter=12, chi2r_target=0.0)
        out = []
        for k in (0, 1):   # forward, adjoint
            cnt, tot = C.c_int(), C.c_double()
            assert ctx.L.gpair_timing_read(ctx.h, k, C.byref(cnt), C.byref(tot)) == 0
            out.append((cnt.value, tot.value))
        assert ctx.L.gpair_timing_enable(ctx.h, 0) == 0
    return res, out


def test_dft_timing_samples_every_nth_launch():
    """the same run timed at every launch and at every 4th: the sampled counts are
    the ceiling of a quarter of the full ones, every duration positive, and the
    loop's result does not change"""
    prob = synth.make_problem("cfg1", seed=1)
    r1, t1 = _timed_run(prob, 1)
    r4, t4 = _timed_run(prob, 4)
    print("every launch", t1, "every 4th", t4, flush=True)
    assert (r1.iterations, r1.forward_dfts, r1.adjoint_dfts) == (r4.iterations, r4.forward_dfts, r4.adjoint_dfts)
    assert r1.J == r4.J
    for (n1, ms1), (n4, ms4) in zip(t1, t4):
        assert n1 > 0 and ms1 > 0.0 and ms4 > 0.0
        assert n4 == math.ceil(n1 / 4)
