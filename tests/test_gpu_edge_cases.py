"""Edge cases of the device solve (needs an MI355X): all-zero sources, the
largest component batch, the smallest tree, a domain with unequal sides and
blocks only along one face (ragged block lists per level), and the argument
errors of the boundary.  Tolerances are those of the parity tests: relative
L2 1e-10 against the FP64 oracle; zero and batch results bit for bit."""
import numpy as np
import pytest

from iblgf_amd import FLAG_LEAF, KERNEL_IF, IblgfError, PoissonSolver, Tree, setups

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _rings():
    return setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                              radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)


def test_zero_sources_give_exact_zeros():
    print("case: all-zero sources, Laplace and IF, after a solve of noise", flush=True)
    v = _rings()
    s = PoissonSolver(v.tree)
    fs, fd = s.field(), s.field()
    rng = np.random.default_rng(3)
    s.upload(fs, rng.standard_normal(s.shape))
    s.apply_lgf(fs, fd)  # leave noise in every work array
    s.upload(fs, np.zeros(s.shape))
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    s.apply_lgf(fs, fd)
    assert not np.any(s.download(fd)[leaves])
    s.apply_lgf_IF(fs, fd, 0.7)
    assert not np.any(s.download(fd))


def test_largest_component_batch_equals_single_solves():
    print("case: 64 components batched (the maximum) vs single solves; 65 refused", flush=True)
    g = setups.UniformGaussians(seed=1000)
    s = PoissonSolver(g.tree)
    rng = np.random.default_rng(4)
    k = 64
    fi = [s.field() for _ in range(k)]
    fo = [s.field() for _ in range(k)]
    for f in fi:
        s.upload(f, rng.standard_normal(s.shape))
    s.apply_lgf_batch(fi, fo)
    got = [s.download(f) for f in fo]
    one = s.field()
    for c in (0, 17, 63):
        s.apply_lgf(fi[c], one)
        assert np.array_equal(s.download(one), got[c]), c
    extra = s.field()
    with pytest.raises(IblgfError):
        s.apply_lgf_batch(fi + [extra], fo + [one])


def test_smallest_tree_matches_oracle(oracle_mod):
    print("case: a single base block of 8^3 cells", flush=True)
    t = Tree((-32,) * 3, (64,) * 3, 8, 1.0 / 32, [(0, 0, 0, 8, 8, 8)])
    t.init_refine(0)
    t.build_masks()
    a, info, n = t.export(), t.info(), t.n
    leaves = (a["flags"] & FLAG_LEAF) != 0
    assert leaves.sum() == 1
    rng = np.random.default_rng(5)
    src = np.zeros((info["nblocks"], n, n, n))
    src[leaves, 1:-1, 1:-1, 1:-1] = rng.standard_normal((1, n - 2, n - 2, n - 2))
    s = PoissonSolver(t)
    fs, fd = s.field(), s.field()
    s.upload(fs, src)
    s.apply_lgf(fs, fd)
    ref = oracle_mod.Tree(a, info, n).apply_lgf(src)
    assert _rel(s.download(fd)[leaves], ref[leaves]) < 1e-10


def test_unequal_domain_sides_and_one_sided_refinement_match_oracle(oracle_mod):
    print("case: base region 6 x 2 x 4 blocks of 8^3, refined only at the +x end", flush=True)
    t = Tree((-64, -32, -48), (128, 64, 96), 8, 1.0 / 32, [(-24, -8, -16, 48, 16, 32)])

    def cond(b, ext, rl, diff):
        return (b[0] + ext) / 2 ** rl > 16

    t.init_refine(2, cond)
    t.build_masks()
    a, info, n = t.export(), t.info(), t.n
    assert info["depth"] > info["base_level"] + 1
    leaves = (a["flags"] & FLAG_LEAF) != 0
    rng = np.random.default_rng(6)
    src = np.zeros((info["nblocks"], n, n, n))
    src[leaves, 1:-1, 1:-1, 1:-1] = rng.standard_normal((int(leaves.sum()), n - 2, n - 2, n - 2))
    s = PoissonSolver(t)
    fs, fd = s.field(), s.field()
    s.upload(fs, src)
    s.apply_lgf(fs, fd)
    ot = oracle_mod.Tree(a, info, n)
    assert _rel(s.download(fd)[leaves], ot.apply_lgf(src)[leaves]) < 1e-10
    s.apply_lgf_IF(fs, fd, 0.4)
    assert _rel(s.download(fd), ot.apply_if(src, 0.4)) < 1e-10


def test_boundary_argument_errors():
    print("case: argument errors of the solve entry points", flush=True)
    g = setups.UniformGaussians(seed=1000)
    s = PoissonSolver(g.tree)
    fs, fd = s.field(), s.field()
    with pytest.raises(IblgfError):
        s.apply_lgf(fs, fs)  # source and target must differ
    with pytest.raises(IblgfError):
        s.apply_lgf(fs, 12345)  # unknown field
    with pytest.raises(ValueError):
        s.upload(fs, np.zeros(10))  # wrong size
    with pytest.raises(IblgfError):
        s.apply_lgf_batch([fs], [fd], kernel=7)
