"""Learned start (gpair_amd/learned.py with the committed network
gpair_amd/weights/unet_cfg5.npz) against the plain start from the prior, on
held-out problems of the network's training family (MI355X only).

The plain MEM-CG run is the judge: a learned run is accepted when it reaches
chi2/ndof <= 1 with an objective within 1e-4 relative of the plain run's or
below it (the learned run goes on past chi2/ndof = 1 until it gets there); a
rejected start costs the learned run plus the plain run.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from gpair_amd import learned, synth
from gpair_amd.api import GpairContext

pytestmark = pytest.mark.gpu

HELD_OUT = range(learned.TEST_SEED0, learned.TEST_SEED0 + 64)


def test_acceptance_rule_falls_back_to_the_plain_run():
    prob = synth.make_problem("cfg1", seed=1)
    alpha = prob.meta["alpha"]
    with GpairContext.from_problem(prob) as ctx:
        # a judge that no start can beat: the learned result is rejected, the judge returned
        judge = SimpleNamespace(J=-1e30)
        out = learned.run_with_start(ctx, prob, alpha, prob.start, None, max_iter=3000, judge=judge, budget=3000)
        assert not out.accepted and out.result is judge and "objective" in out.reason
        # without a judge the fallback is a fresh plain run when the start cannot converge
        out = learned.run_with_start(ctx, prob, alpha, prob.start, None, max_iter=2, judge=None)
        assert not out.accepted and out.plain_result is out.result and out.result.iterations >= 1


def test_learned_start_against_the_plain_start():
    """What the network gains on 64 held-out skies: per problem the plain run from the
    prior (the judge) against the learned path (inference, the run from the learned
    start to the judge's objective, plus the plain run when the start is rejected).
    Every accepted start meets the acceptance rule, and the learned path needs fewer
    iterations than the plain start at the median over the problems the plain start
    solves (iterations: deterministic; the wall-clock medians are printed)."""
    net = learned.UNet.load(learned.WEIGHTS)
    assert net.nparam > 1_900_000
    rows = learned.evaluate(net, (p for _, p in learned.problems_ahead(HELD_OUT)))
    for r in rows:
        if r["accepted"]:
            assert r["learned_conv"] and r["learned_J"] <= r["plain_J"] + 1e-4 * abs(r["plain_J"])
        else:
            assert r["reason"]
    s = learned.summarize(rows)
    print(f"held-out: {s}", flush=True)
    assert s["plain_converged"] >= 32
    assert s["median_path_its"] < s["median_plain_its"]
