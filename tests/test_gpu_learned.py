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


def test_learned_start_gains_on_held_out_problems():
    net = learned.UNet.load(learned.WEIGHTS)
    probs = [learned.family_problem(s) for s in HELD_OUT] + [synth.make_problem("cfg2", seed=1)]
    rows = learned.evaluate(net, None, problems=probs)
    plain = sum(r["plain_s"] for r in rows)
    # cost of the learned path: inference + learned run, plus the plain run when rejected
    cost = sum(r["learned_s"] + (0.0 if r["accepted"] else r["plain_s"]) for r in rows)
    its_ratio = np.median([r["learned_its"] / r["plain_its"] for r in rows])
    n_acc = sum(r["accepted"] for r in rows)
    print(f"plain {plain:.2f} s, learned path {cost:.2f} s, median iteration ratio {its_ratio:.2f}, "
          f"accepted {n_acc}/{len(rows)}", flush=True)
    assert all(r["plain_conv"] for r in rows)
    assert n_acc >= len(rows) - 1
    assert cost < 0.85 * plain
    assert its_ratio < 0.85
