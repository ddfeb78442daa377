"""The oracle's integrating-factor solve (PoissonSolver::apply_lgf_IF,
include/iblgf/solver/poisson/poisson.hpp:427-494) pinned without the
reference:

* single level: it equals the direct separable convolution with
  E_alpha(x) = prod_i e^{-2 alpha} I_{x_i}(2 alpha) (lgf_ge.hpp:158-168), with
  scipy's ``ive`` as the Bessel routine, over the whole mesh;
* several levels: e^{alpha_l Lap_h} with alpha_l = alpha_base 4^(l - base)
  (lgf_ge.hpp:83) is the heat semigroup at time t = alpha_base dx_base^2 on
  every level, so a Gaussian source comes back as the closed-form heat
  solution on the leaves of each level, to the discretisation error.  A level
  scale off by a factor 4 changes the finest level's result by a quarter of
  the diffusion, far outside the tolerance.
"""
import numpy as np
import pytest
from scipy.ndimage import convolve1d
from scipy.special import ive

from iblgf_amd import FLAG_CORRECTION, FLAG_LEAF, Tree, setups
from oracle import oracle as O


def test_single_level_if_equals_direct_convolution():
    g = setups.UniformGaussians(seed=1000)
    src, a, n = g.source(), g.arrays, g.tree.n
    alpha = 0.5
    phi = O.Tree(a, g.info, n).apply_if(src, alpha)
    off, half = 12, g.half
    dense = np.zeros((g.ext + 2 * off,) * 3)
    for s in np.nonzero(a["flags"] & FLAG_LEAF)[0]:
        b = a["bbase"][s] + half + off
        e = g.block_extent
        dense[b[2]:b[2] + e, b[1]:b[1] + e, b[0]:b[0] + e] = src[s, 1:-1, 1:-1, 1:-1]
    k = ive(np.arange(-off, off + 1), 2 * alpha)
    out = dense
    for ax in range(3):
        out = convolve1d(out, k, axis=ax, mode="constant")
    err = 0.0
    for s in range(g.info["nblocks"]):
        if a["level"][s] != g.info["base_level"]:
            continue
        b = a["bbase"][s] + half + off - 1  # every point of the block, buffers included
        d = out[b[2]:b[2] + n, b[1]:b[1] + n, b[0]:b[0] + n]
        err = max(err, np.abs(phi[s] - d).max())
    rel = err / np.abs(out).max()
    print(f"single-level IF vs direct convolution: rel max {rel:.3e}")
    assert rel < 1e-6  # kernel truncated to the 27 neighbour blocks: E(9) ~ 1e-8


@pytest.mark.parametrize("alpha_base", [0.5, 1.5])
def test_multilevel_if_is_the_heat_semigroup(alpha_base):
    e, dxb, sig2 = 8, 1.0 / 32, 0.08 ** 2
    tree = Tree((-48,) * 3, (96,) * 3, e, dxb, [(-24, -24, -24, 48, 48, 48)])

    def refine(base, ext, rl, diff):
        dx = dxb / 2 ** rl
        lo = (np.array(base) - 2 + 0.5) * dx
        hi = (np.array(base) + ext + 2 - 0.5) * dx
        return bool(np.linalg.norm(np.clip(0.0, lo, hi)) < 0.12)

    tree.init_refine(1, refine)
    tree.build_masks()
    a, info, n = tree.export(), tree.info(), tree.n
    src = np.zeros((info["nblocks"], n, n, n))
    leaves = np.nonzero(a["flags"] & FLAG_LEAF)[0]
    pts = {s: setups.cell_centres(a, s, n, dxb, info["base_level"]) for s in leaves}
    for s in leaves:
        src[s][1:-1, 1:-1, 1:-1] = np.exp(-(pts[s] ** 2).sum(-1) / (2 * sig2))[1:-1, 1:-1, 1:-1]
    phi = O.Tree(a, info, n).apply_if(src, alpha_base)
    t = alpha_base * dxb ** 2
    amp = (sig2 / (sig2 + 2 * t)) ** 1.5
    errs, change = {}, 0.0
    for s in leaves:
        if a["flags"][s] & FLAG_CORRECTION:
            continue
        r2 = (pts[s] ** 2).sum(-1)[1:-1, 1:-1, 1:-1]
        u = amp * np.exp(-r2 / (2 * (sig2 + 2 * t)))
        change = max(change, np.abs(u - np.exp(-r2 / (2 * sig2))).max())
        L = int(a["level"][s])
        errs[L] = max(errs.get(L, 0.0), np.abs(phi[s][1:-1, 1:-1, 1:-1] - u).max())
    print(f"alpha_base {alpha_base}: max diffusion {change:.4f}, error by level {errs}")
    assert sorted(errs) == [info["base_level"], info["base_level"] + 1]
    for L, err in errs.items():
        assert err < 0.03 * change
