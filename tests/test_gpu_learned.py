"""Learned start (gpair_amd/learned.py with the committed network
gpair_amd/weights/unet_cfg5.npz) against the plain start from the prior
(MI355X only).

The plain MEM-CG run is the judge: a learned run is accepted when it reaches
chi2/ndof <= 1 with an objective no worse than the plain run's by more than
1e-4 relative; a rejected start costs the learned run plus the plain run.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from gpair_amd import learned, synth
from gpair_amd.api import GpairContext

pytestmark = pytest.mark.gpu

HELD_OUT = [900000, 900001, 900002, 900003]


def test_acceptance_rule_falls_back_to_the_plain_run():
    prob = synth.make_problem("cfg1", seed=1)
    alpha = prob.meta["alpha"]
    with GpairContext.from_problem(prob) as ctx:
        # a judge that no start can beat: the learned result is rejected, the judge returned
        judge = SimpleNamespace(J=-1e30)
        out = learned.run_with_start(ctx, prob, alpha, prob.start, None, max_iter=3000, judge=judge)
        assert not out.accepted and out.result is judge and "objective" in out.reason
        # without a judge the fallback is a fresh plain run when the start cannot converge
        out = learned.run_with_start(ctx, prob, alpha, prob.start, None, max_iter=2, judge=None)
        assert not out.accepted and out.plain_result is out.result and out.result.iterations >= 1


def test_learned_start_against_the_plain_start():
    """what the network gains: per problem, the plain run from the prior (the judge)
    against the learned start (inference + learned run, plus the plain run when the
    start is rejected).  Every accepted start meets the acceptance rule; on the
    held-out problems of the training family most starts are accepted and the
    learned path costs no more than the plain one; the headline problem's outcome
    is printed (its time to chi2/ndof = 1 varies with the trajectory: 2,760 to
    4,100 iterations from the learned start, 5,200 to 6,000 from the prior)."""
    net = learned.UNet.load(learned.WEIGHTS)
    probs = [learned.family_problem(s) for s in HELD_OUT] + [synth.make_problem("cfg2", seed=1)]
    rows = learned.evaluate(net, None, problems=probs)
    for r in rows:
        if r["accepted"]:
            assert r["learned_conv"] and r["learned_J"] <= r["plain_J"] + 1e-4 * abs(r["plain_J"])
        else:
            assert r["reason"]
    fam = [r for r in rows if r["cfg"] == "cfg5" and r["plain_conv"]]
    plain = sum(r["plain_s"] for r in fam)
    cost = sum(r["learned_s"] for r in fam)
    dJ = [1.0 - r["learned_J"] / r["plain_J"] for r in fam if r["accepted"]]
    head = rows[-1]
    print(f"family: {len(fam)} problems, plain {plain:.2f} s, learned path {cost:.2f} s, accepted "
          f"{sum(r['accepted'] for r in fam)}, objective lower by {np.round(dJ, 4).tolist()}; headline: plain "
          f"{head['plain_its']} iterations, learned {head['learned_its']} accepted {head['accepted']}", flush=True)
    assert len(fam) >= 3
    assert sum(r["accepted"] for r in fam) >= len(fam) - 1
    assert cost <= 1.1 * plain
