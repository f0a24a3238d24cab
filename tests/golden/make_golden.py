"""Writes golden vectors from the reference's own CPU model (gpair1/trunk/cpu.c,
compiled by `make -C oracle ref` into oracle/_ref/libgpair_refcpu.so).

    python tests/golden/make_golden.py

Each fixture holds its inputs (uv, packed data, T3 references, image) and the
reference's float outputs: visibilities (image2vis), mock data (vis2data),
chi2 (data2chi2), data gradient (compute_data_gradient), entropy and entropy
gradient (GullSkilling_entropy[_gradient]), and set_model priors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from gpair_amd import synth  # noqa: E402
from oracle import RefCPU, ref_available  # noqa: E402


def quiet(fn, *a):
    """the reference prints a line per image2vis call"""
    sys.stdout.flush()
    fd = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(fd, 1)
    try:
        return fn(*a)
    finally:
        os.dup2(saved, 1)
        os.close(fd)


def one(name, prob, img_seed):
    R = RefCPU(prob)
    rng = np.random.default_rng(img_seed)
    img = (prob.start * (0.5 + rng.random(prob.W * prob.W))).astype(np.float32)
    chi2, vis, mock = quiet(R.image2chi2, img)
    grad = quiet(R.data_gradient, img)
    ent = R.entropy(img, prob.model)
    eg = R.entropy_grad(img, prob.model)
    priors = np.stack([R.set_model(prob.pixellation, t, 1.0, 1.0) for t in range(5)])
    np.savez_compressed(
        os.path.join(HERE, f"{name}.npz"), W=prob.W, pix=np.float32(prob.pixellation), u=prob.u, v=prob.v,
        npow=prob.npow, nbis=prob.nbis, data=prob.data, err=prob.err, phasor=prob.phasor, uvpnt=prob.uvpnt,
        sign=prob.sign, model=prob.model, image=img, ref_chi2=np.float32(chi2), ref_vis=vis, ref_mock=mock,
        ref_grad=grad, ref_entropy=np.float32(ent), ref_entropy_grad=eg, ref_priors=priors)
    print(f"{name}: W={prob.W} nuv={prob.nuv} nbis={prob.nbis} chi2={chi2:.6g} S={ent:.6g}")


def main():
    if not ref_available():
        raise SystemExit("oracle/_ref/libgpair_refcpu.so is not built (make -C oracle ref)")
    one("ref_cpu_w64", synth.make_problem("cfg1", seed=5), 1)
    one("ref_cpu_w64_7tel", synth.make_problem("cfg1", seed=9, ntel=7, nch=3, nsnap=2), 2)


if __name__ == "__main__":
    main()
