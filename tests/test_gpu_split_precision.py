"""Split-precision solves: every 1-D DFT leg a GEMM on the matrix cores
(BF16 three-term split, FP16 three-term and one-term splits with power-of-two
line scaling, four signed 7-bit integer digits with exact INT32 sums), the
spectra and the multiply-accumulate as on the mixed path, fields and level
sums FP64.  Each is compared with the FP64 oracle per field component.

Tolerances (relative L2 over the leaves, per component), set from the
arithmetic of each variant: a BF16 hi/lo pair carries 16 significant bits and
FP32 accumulation follows, so the legs err at ~1e-5 (bound 2e-4); FP16 one
term carries 11 bits (~1e-3, bound 1e-2); the FP16 hi/lo pair of a scaled line
carries 22 bits relative to the line's maximum (bound 5e-6); the integer
digits carry 27 bits and sum exactly, so only the FP32 storage of leg outputs
remains (bound 2e-6).  With FP16 three-term legs in the forward transforms
only (the inverse legs FP64, as the mixed path) the solve meets the 1e-6
target (bound 1e-6).  Batched solves equal single solves bit for bit (same
kernels, same order), and the solves are linear to FP32 storage precision.
"""
import numpy as np
import pytest

from iblgf_amd import (FLAG_LEAF, PREC_BF16X3, PREC_FP16, PREC_FP16X3, PREC_FP16X3_FORWARD, PREC_INT8X4,
                       PoissonSolver, setups)

pytestmark = pytest.mark.gpu

BOUND = {PREC_BF16X3: 2e-4, PREC_FP16X3: 5e-6, PREC_FP16: 1e-2, PREC_INT8X4: 2e-6, PREC_FP16X3_FORWARD: 1e-6}
NAME = {PREC_BF16X3: "bf16x3", PREC_FP16X3: "fp16x3", PREC_FP16: "fp16", PREC_INT8X4: "int8x4",
        PREC_FP16X3_FORWARD: "fp16x3 forward"}


def _rel(a, b):
    d = a - b
    return np.linalg.norm(d) / np.linalg.norm(b), np.max(np.abs(d)) / np.max(np.abs(b))


def _small_rings(**kw):
    base = dict(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0, radius=(4.0, 8.0),
                core=(1.5, 2.5), refinement_factor=0.5)
    base.update(kw)
    return setups.VortexRings(**base)


_CACHE = {}


def _case(oracle_mod, extent):
    if extent not in _CACHE:
        v = _small_rings(block_extent=extent)
        srcs = [v.source(c) for c in range(3)]
        ot = oracle_mod.Tree(v.arrays, v.info, v.tree.n)
        _CACHE[extent] = (v, srcs, [ot.apply_lgf(s) for s in srcs])
    return _CACHE[extent]


@pytest.mark.parametrize("prec", [PREC_INT8X4, PREC_FP16X3, PREC_BF16X3, PREC_FP16, PREC_FP16X3_FORWARD])
@pytest.mark.parametrize("extent", [8, 16])
def test_split_precision_error_per_component(oracle_mod, extent, prec):
    print(f"case: {NAME[prec]} legs, adaptive rings, block extent {extent}", flush=True)
    v, srcs, refs = _case(oracle_mod, extent)
    solver = PoissonSolver(v.tree)
    solver.set_precision(prec)
    fin = [solver.field() for _ in srcs]
    fout = [solver.field() for _ in srcs]
    for f, x in zip(fin, srcs):
        solver.upload(f, x)
    solver.apply_lgf_batch(fin, fout)
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    got = [solver.download(f) for f in fout]
    for c in range(3):
        l2, mx = _rel(got[c][leaves], refs[c][leaves])
        print(f"{NAME[prec]} n={extent + 2} component {c}: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
        assert l2 < BOUND[prec]
    # one component alone: bitwise the batched result
    solver.apply_lgf(fin[1], fout[0])
    assert np.array_equal(solver.download(fout[0]), got[1])


def test_uniform_gaussians_int8_digits_match_direct_sum(oracle_mod):
    """The single-level case against the oracle, whose result the FP64 path
    matches with the brute-force direct LGF sum."""
    g = setups.UniformGaussians(seed=1000)
    src = g.source()
    solver = PoissonSolver(g.tree)
    solver.set_precision(PREC_INT8X4)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.apply_lgf(fs, fd)
    ref = oracle_mod.Tree(g.arrays, g.info, g.tree.n).apply_lgf(src)
    leaves = (g.arrays["flags"] & FLAG_LEAF) != 0
    l2, mx = _rel(solver.download(fd)[leaves], ref[leaves])
    print(f"uniform int8x4: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
    assert l2 < BOUND[PREC_INT8X4]


@pytest.mark.parametrize("prec", [PREC_INT8X4, PREC_BF16X3])
def test_split_precision_is_linear(prec):
    """a x + b y solves to a phi(x) + b phi(y) within the variant's error."""
    v, srcs, _ = _CACHE.get(16) or (None, None, None)
    if v is None:
        v = _small_rings(block_extent=16)
        srcs = [v.source(c) for c in range(3)]
    solver = PoissonSolver(v.tree)
    solver.set_precision(prec)
    fs, fd = solver.field(), solver.field()
    outs = []
    for x in (srcs[0], srcs[2], 0.5 * srcs[0] - 2.0 * srcs[2]):
        solver.upload(fs, x)
        solver.apply_lgf(fs, fd)
        outs.append(solver.download(fd))
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    comb = 0.5 * outs[0] - 2.0 * outs[1]
    l2, _ = _rel(outs[2][leaves], comb[leaves])
    print(f"{NAME[prec]} linearity: relL2={l2:.3e}", flush=True)
    assert l2 < 2 * BOUND[prec]
