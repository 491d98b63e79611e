"""Shared pieces of the differentiable-render tests: FP32 parameters around
the workload's fp16 weights, the L2 loss, and central finite differences of
the FP64 oracle forward."""
import numpy as np

DIFF_W = DIFF_H = 24
FD_EPS = 1e-3


def params32(wl, seed=11, sigma=0.01):
    """The workload's fp16 parameter blocks as float32, perturbed by N(0, sigma)."""
    rng = np.random.default_rng(seed)
    out = []
    for p in wl.params:
        v = np.asarray(p, np.uint16).view(np.float16).astype(np.float32)
        out.append((v + rng.normal(0.0, sigma, v.shape)).astype(np.float32))
    return out


def target_image(n, seed=12):
    return np.random.default_rng(seed).uniform(0.0, 1.0, (n, 3))


def loss(rgb, target):
    return float(np.sum((np.asarray(rgb, np.float64) - target) ** 2))


def pick_params(grads, seed=13, per_block=(16, 24, 24)):
    """64 (block, index) pairs among the larger oracle gradients of each block."""
    rng = np.random.default_rng(seed)
    picks = []
    for b, (g, k) in enumerate(zip(grads, per_block)):
        nz = np.flatnonzero(g)
        order = nz[np.argsort(-np.abs(g[nz]))]
        top = order[: max(4 * k, k)]
        for i in rng.choice(top, k, replace=False):
            picks.append((b, int(i)))
    return picks


def fd_grads(model, rm, pix, params, target, picks, eps=FD_EPS):
    """Central differences of the L2 loss of the FP64 oracle forward.  The step
    actually taken is the float32 difference of the perturbed parameter."""
    out = []
    for b, i in picks:
        p = [x.copy() for x in params]
        v = params[b][i]
        p[b][i] = np.float32(v + eps)
        hp = float(p[b][i]) - float(v)
        lp = loss(model.diff_render(rm, pix, p)[0], target)
        p[b][i] = np.float32(v - eps)
        hm = float(v) - float(p[b][i])
        lm = loss(model.diff_render(rm, pix, p)[0], target)
        out.append((lp - lm) / (hp + hm))
    return np.array(out)
