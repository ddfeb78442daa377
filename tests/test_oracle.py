"""The CPU oracle pinned by the reference's known-answer tests and by brute
force: block convolution (tests/conv_test/convSingleBlockTest.cpp:60-106,
convFFTTest.cpp:122-225, restated at 1e-12), the single-level solve against
the direct LGF sum, and the faithful (per-pair transform) and cached modes."""
import numpy as np
import pytest

from iblgf_amd import FLAG_LEAF, setups
from oracle import oracle as O


def test_delta_source_reproduces_lgf_values():
    # convSingleBlockTest: a unit source at the block origin convolved with the
    # LGF gives lgf.get(offset) on the target block.
    n = 10
    src = np.zeros((n, n, n))
    src[0, 0, 0] = 1.0
    for shift in [(0, 0, 0), (8, 0, 0), (-16, 8, 24)]:
        out = O.block_conv(src, shift)
        for (k, j, i) in [(0, 0, 0), (3, 5, 7), (9, 9, 9), (0, 9, 4)]:
            ref = O.lgf_get(shift[0] + i, shift[1] + j, shift[2] + k)
            assert abs(out[k, j, i] - ref) < 1e-12 * max(abs(ref), 1e-3)


@pytest.mark.parametrize("step", [1, 2, 8])
def test_fft_convolution_equals_direct_sum(step):
    # convFFTTest: the zero-padded FFT convolution equals the direct sum.
    n = 8
    rng = np.random.default_rng(3 + step)
    src = rng.standard_normal((n, n, n))
    shift = (16, -8, 0)
    out = O.block_conv(src, shift, step=step, scale=0.25)
    kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    sx = np.stack([ii, jj, kk], -1).reshape(-1, 3)
    for t in [(0, 0, 0), (7, 7, 7), (3, 1, 6)]:
        tx = np.array([[shift[0] + t[0], shift[1] + t[1], shift[2] + t[2]]])
        # G sampled at step * (x_t - x_s)
        d = (tx - sx) * step
        ref = 0.25 * sum(O.lgf_get(*row) * q for row, q in zip(d, src.reshape(-1)))
        assert abs(out[t[2], t[1], t[0]] - ref) < 1e-12 * np.abs(out).max()


def test_single_level_solve_equals_direct_sum():
    g = setups.UniformGaussians(seed=1000)
    src = g.source()
    a, n = g.arrays, g.tree.n
    ot = O.Tree(a, g.info, n)
    phi = ot.apply_lgf(src)
    leaves = np.nonzero(a["flags"] & FLAG_LEAF)[0]
    sx, q = [], []
    for s in leaves:
        b = a["bbase"][s]
        kk, jj, ii = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), np.arange(1, n - 1), indexing="ij")
        sx.append(np.stack([b[0] + ii - 1, b[1] + jj - 1, b[2] + kk - 1], -1).reshape(-1, 3))
        q.append(src[s, 1:-1, 1:-1, 1:-1].reshape(-1))
    sx, q = np.concatenate(sx), np.concatenate(q)
    for s in leaves[[0, 21, 63]]:
        b = a["bbase"][s]
        kk, jj, ii = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        tx = np.stack([b[0] + ii - 1, b[1] + jj - 1, b[2] + kk - 1], -1).reshape(-1, 3)
        direct = O.direct_sum(sx, q, tx, g.info["dx_base"] ** 2)
        assert np.max(np.abs(phi[s].reshape(-1) - direct)) < 1e-12 * np.max(np.abs(direct))


def test_faithful_and_cached_modes_agree():
    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    src = v.source(1)
    ot = O.Tree(v.arrays, v.info, v.tree.n)
    a = ot.apply_lgf(src, faithful=False)
    b = ot.apply_lgf(src, faithful=True)
    assert np.array_equal(a, b)


def test_multilevel_solve_reduces_to_leaf_laplacian():
    """On the finest leaves, the discrete Laplacian of the solution returns the
    source up to the interpolation error of the level coupling (the check of
    PoissonSolver::apply_laplace, include/iblgf/solver/poisson/poisson.hpp:1290-1360)."""
    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    src = v.source(2)
    a, n, info = v.arrays, v.tree.n, v.info
    phi = O.Tree(a, info, n).apply_lgf(src)
    finest = info["depth"] - 1
    dx = info["dx_base"] / 2 ** (finest - info["base_level"])
    err, ref = 0.0, 0.0
    for s in np.nonzero((a["flags"] & FLAG_LEAF) & (a["level"] == finest))[0]:
        p = phi[s]
        lap = (p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1] + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1]
               + p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2] - 6 * p[1:-1, 1:-1, 1:-1]) / dx ** 2
        err = max(err, np.max(np.abs(lap - src[s, 1:-1, 1:-1, 1:-1])))
        ref = max(ref, np.max(np.abs(src[s])))
    assert err < 0.05 * ref


@pytest.mark.parametrize("ranks", [0, 1, 3])
def test_rank_split_baseline_equals_single_solve(ranks):
    """The CPU baseline's warm solves (kernel spectra kept between solves,
    targets split over workers as contiguous load ranges) give the single
    solve's result bit for bit: each target sums its sources in list order
    whatever worker runs it, and every solve starts from cleared work fields."""
    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    src = v.source(0)
    ot = O.Tree(v.arrays, v.info, v.tree.n)
    ref = ot.apply_lgf(src)
    sec, got = ot.lgf_baseline(src, ranks=ranks, nsolves=2)
    assert len(sec) == 2 and np.all(sec > 0)
    assert np.array_equal(got, ref)


def test_stream_solve_is_the_base_level_direct_sum():
    """fmm_type STREAM (poisson.hpp:609-611, 645-647; masks fmm.hpp:199-233):
    one FMM pass at the base level from the interiors of the non-correction
    base-level blocks (base blocks missing a neighbour are correction blocks,
    domain.hpp:617-735), written on the base-level leaves, equals the exact
    all-pairs LGF sum over those cells (the base level as a uniform mesh),
    whatever the finer levels hold.  The STREAM masks follow the rule:
    base-level blocks targets, non-correction base-level blocks sources,
    parents marked upward, nothing finer."""
    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    a, n, info = v.arrays, v.tree.n, v.info
    bl = info["base_level"]
    one = np.uint64(1)
    ms, mt = a["mask_src"] & one, a["mask_tgt"] & one
    lvl = a["level"]
    corr = (a["flags"] & 2) != 0
    assert np.array_equal(mt[lvl == bl], np.ones(np.sum(lvl == bl), np.uint64))
    assert np.array_equal(ms[lvl == bl] == 1, ~corr[lvl == bl])
    assert not ms[lvl > bl].any() and not mt[lvl > bl].any()
    for s in np.nonzero(lvl < bl)[0]:
        kids = [c for c in a["children"][s] if c >= 0]
        assert ms[s] == (1 if any(ms[c] for c in kids) else 0)
    rng = np.random.default_rng(3)
    src = rng.standard_normal((info["nblocks"], n, n, n))
    phi = O.Tree(a, info, n).apply_lgf_masked(src, 0)
    base = np.nonzero(lvl == bl)[0]
    sx, q = [], []
    kk, jj, ii = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), np.arange(1, n - 1), indexing="ij")
    for s in base:
        if corr[s]:
            continue
        b = a["bbase"][s]
        sx.append(np.stack([b[0] + ii - 1, b[1] + jj - 1, b[2] + kk - 1], -1).reshape(-1, 3))
        q.append(src[s, 1:-1, 1:-1, 1:-1].reshape(-1))
    sx, q = np.concatenate(sx), np.concatenate(q)
    leaves = [s for s in base if a["flags"][s] & FLAG_LEAF]
    assert leaves
    K, J, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    for s in leaves[:3]:
        b = a["bbase"][s]
        tx = np.stack([b[0] + I - 1, b[1] + J - 1, b[2] + K - 1], -1).reshape(-1, 3)
        direct = O.direct_sum(sx, q, tx, info["dx_base"] ** 2)
        assert np.max(np.abs(phi[s].reshape(-1) - direct)) < 1e-11 * np.max(np.abs(direct))
    # nothing outside the base-level leaves is written
    other = np.ones(info["nblocks"], bool)
    other[leaves] = False
    assert not phi[other].any()


def test_ib2amr_single_level_is_the_direct_sum_from_ib_blocks():
    """fmm_type IB2AMR (pressure correction, linsys.hpp:131-145; masks
    fmm.hpp:146-195) on a one-level mesh: the blocks that hold
    immersed-boundary points are the only sources and every block a target,
    so the solve on the leaves equals the direct LGF sum over the interiors of
    the is_ib blocks.  The masks follow the rule per level (mask index 4 + rl)."""
    from iblgf_amd import Tree, setups as S

    # 4^3 base blocks: every pair of blocks is in each other's neighbour or
    # influence lists, so the pass is exact (no far field through coarser levels)
    t = Tree((-16, -16, -16), (32, 32, 32), 8, 1.0 / 32, [(-16, -16, -16, 32, 32, 32)])
    pts = S.fibonacci_sphere(40, 0.03) - 0.3  # near a corner: only some blocks hold (or border) points
    t.set_ib(pts, 0, 2, 2.0)
    t.init_refine(0)
    t.build_masks()
    a, info, n = t.export(), t.info(), t.n
    isib = (a["flags"] & 16) != 0
    bl = info["base_level"]
    one = np.uint64(1)
    ms, mt = (a["mask_src"] >> np.uint64(4)) & one, (a["mask_tgt"] >> np.uint64(4)) & one
    base = a["level"] == bl
    assert isib[base].any() and not isib[base].all()
    assert np.array_equal(ms[base] == 1, isib[base]) and np.all(mt[base] == 1)
    rng = np.random.default_rng(5)
    src = rng.standard_normal((info["nblocks"], n, n, n))
    phi = O.Tree(a, info, n).apply_lgf_masked(src, 4)
    sx, q = [], []
    kk, jj, ii = np.meshgrid(np.arange(1, n - 1), np.arange(1, n - 1), np.arange(1, n - 1), indexing="ij")
    for s in np.nonzero(base & isib)[0]:
        b = a["bbase"][s]
        sx.append(np.stack([b[0] + ii - 1, b[1] + jj - 1, b[2] + kk - 1], -1).reshape(-1, 3))
        q.append(src[s, 1:-1, 1:-1, 1:-1].reshape(-1))
    sx, q = np.concatenate(sx), np.concatenate(q)
    K, J, I = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    leaves = np.nonzero(base & ((a["flags"] & FLAG_LEAF) != 0))[0]
    for s in leaves[[0, len(leaves) // 2, -1]]:
        b = a["bbase"][s]
        tx = np.stack([b[0] + I - 1, b[1] + J - 1, b[2] + K - 1], -1).reshape(-1, 3)
        direct = O.direct_sum(sx, q, tx, info["dx_base"] ** 2)
        assert np.max(np.abs(phi[s].reshape(-1) - direct)) < 1e-11 * np.max(np.abs(direct))


def test_ib2amr_masks_mark_ib_ancestors_per_level():
    s = setups.IbSphere(npoints=600, block_extent=8, nref=2, half_width=1.0)
    a, info = s.arrays if hasattr(s, "arrays") else s.tree.export(), s.tree.info()
    bl, depth = info["base_level"], info["depth"]
    isib = (a["flags"] & 16) != 0
    sub = isib.copy()  # is_ib or a descendant that is
    for l in range(depth - 2, bl - 1, -1):
        for o in np.nonzero(a["level"] == l)[0]:
            if any(c >= 0 and sub[c] for c in a["children"][o]):
                sub[o] = True
    one = np.uint64(1)
    for l in range(bl, depth):
        idx = np.uint64(4 + l - bl)
        ms, mt = (a["mask_src"] >> idx) & one, (a["mask_tgt"] >> idx) & one
        at = a["level"] == l
        assert np.all(mt[at] == 1) and np.array_equal(ms[at] == 1, sub[at])
        assert not ms[a["level"] > l].any() and not mt[a["level"] > l].any()


def _mesh_points(n, mesh, comp, fine, child):
    """Sample positions of the n^3 points of a parent block (fine False) or of
    its child `child` (fine True) in parent-cell units for a field of mesh type
    mesh and field index comp: cell centres on axes of relative position 1,
    cell nodes on axes of relative position 0.  Face fields (normal comp) sit
    on nodes along comp, edge fields (along comp) on nodes across comp."""
    idx = np.arange(n) - 1.0
    axes = []
    for a in range(3):
        node = (mesh == 1 and a == comp) or (mesh == 2 and a != comp)
        if fine:
            b = (child >> a) & 1
            x = (b * (n - 2) + idx + (0.0 if node else 0.5)) / 2.0
        else:
            x = idx + (0.0 if node else 0.5)
        axes.append(x)
    Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return X, Y, Z


@pytest.mark.parametrize("mesh,comp", [(0, 0), (1, 0), (1, 1), (1, 2), (2, 0), (2, 1), (2, 2)])
def test_mesh_transfers_reproduce_polynomials_at_the_mesh_positions(mesh, comp):
    """The mesh-type level transfers (cell_center_nli_intrp.hpp:599-618) are
    pinned by geometry: the interpolation reproduces a quadratic sampled at the
    field's own positions on the parent, at the child's interior positions; the
    restriction of all eight children reproduces a linear function at the
    parent's interior positions.  The same transfers with any other mesh type
    or field index miss by a quarter of a fine cell times the gradient."""
    n = 18
    f = lambda X, Y, Z: 0.3 + 0.7 * X - 0.4 * Y + 0.9 * Z + 0.05 * X * Y - 0.03 * Z * Z + 0.02 * X * X
    lin = lambda X, Y, Z: 0.3 + 0.7 * X - 0.4 * Y + 0.9 * Z
    inner = (slice(1, -1),) * 3
    # every other placement (a cell mesh has no field index)
    wrong = [(m, c) for m, c in [(0, 0)] + [(m, c) for m in (1, 2) for c in range(3)]
             if (m, c) != (mesh, comp) and not (mesh == 0 and m == 0)]
    for child in range(8):
        parent = f(*_mesh_points(n, mesh, comp, False, 0))
        got = O.mesh_transfer(parent, mesh, comp, 1, child)
        ref = f(*_mesh_points(n, mesh, comp, True, child))
        assert np.max(np.abs(got[inner] - ref[inner])) < 1e-11, child
        for m, c in wrong:
            bad = O.mesh_transfer(parent, m, c, 1, child)
            assert np.max(np.abs(bad[inner] - ref[inner])) > 0.05, (m, c, child)
    out = np.zeros((n, n, n))
    for child in range(8):
        fine = lin(*_mesh_points(n, mesh, comp, True, child))
        O.mesh_transfer(fine, mesh, comp, 0, child, out)
    ref = lin(*_mesh_points(n, mesh, comp, False, 0))
    assert np.max(np.abs(out[inner] - ref[inner])) < 1e-12
    for m, c in wrong:
        out = np.zeros((n, n, n))
        for child in range(8):
            O.mesh_transfer(lin(*_mesh_points(n, mesh, comp, True, child)), m, c, 0, child, out)
        assert np.max(np.abs(out[inner] - ref[inner])) > 0.05, (m, c)
