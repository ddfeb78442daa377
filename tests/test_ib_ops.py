"""CPU checks of the block operators and the immersed-boundary transfer.

The oracle's operators are pinned by discrete identities that a wrong index,
sign or offset breaks: div(grad) equals the 7-point Laplacian, curl(grad) and
div(curl^T) vanish, the Laplacian of a quadratic is its constant.  The delta
kernels are pinned by their moment conditions and the smearing by the
centroid of the spread weights (the face offset).  The host-built delta
weights that the device uses are compared bit for bit with the oracle's
literal loops of the reference (influence blocks, then nodes).
"""
import numpy as np
import pytest

from iblgf_amd import setups


def _poly_field(arrays, n, dx, coef, level):
    """f(x,y,z) = polynomial in the global node index of every block at `level` (buffers included)."""
    nb = arrays["bbase"].shape[0]
    out = np.zeros((nb, n, n, n))
    idx = np.arange(n) - 1
    for s in np.nonzero(arrays["level"] == level)[0]:
        b = arrays["bbase"][s]
        Z, Y, X = np.meshgrid(b[2] + idx, b[1] + idx, b[0] + idx, indexing="ij")
        out[s] = coef(X.astype(float), Y.astype(float), Z.astype(float))
    return out


@pytest.fixture(scope="module")
def uni(oracle_mod):
    g = setups.UniformGaussians(seed=1000)
    return g, oracle_mod.Tree(g.arrays, g.info, g.tree.n)


def _interior(f, arrays=None, level=None):
    if arrays is not None:
        f = f[arrays["level"] == level]
    return f[..., 1:-1, 1:-1, 1:-1]


def test_laplacian_of_quadratic_is_constant(uni):
    g, ot = uni
    n, L = g.tree.n, g.info["base_level"]
    dx = g.info["dx_base"]
    phi = _poly_field(g.arrays, n, dx, lambda X, Y, Z: 3 * X * X - Y * Y + 0.5 * Z * Z + X * Y + 2 * Z, L)
    lap = ot.block_op(0, phi, L)
    want = (6.0 - 2.0 + 1.0) / (dx * dx)
    assert np.allclose(_interior(lap, g.arrays, L), want, rtol=1e-12, atol=0)


def test_div_grad_is_laplacian_and_curl_grad_vanishes(uni):
    g, ot = uni
    n, L = g.tree.n, g.info["base_level"]
    rng = np.random.default_rng(3)
    phi = rng.standard_normal((g.info["nblocks"], n, n, n))
    grad = ot.block_op(1, phi, L)
    # div at i reads grad at i+1, written only for i+1 <= n-2
    div = ot.block_op(2, grad, L)
    lap = ot.block_op(0, phi, L)
    sl = (slice(None), slice(1, -2), slice(1, -2), slice(1, -2))
    assert np.allclose(div[sl], lap[sl], rtol=0, atol=1e-9 * np.abs(lap[sl]).max())
    # curl at i reads grad at i-1, written only for i-1 >= 1
    curl = ot.block_op(3, grad, L)
    sl2 = (slice(None), slice(2, -1), slice(2, -1), slice(2, -1))
    for c in range(3):
        assert np.abs(curl[c][sl2]).max() < 1e-9 * np.abs(grad[0]).max()


def test_div_of_curl_transpose_vanishes(uni):
    g, ot = uni
    n, L = g.tree.n, g.info["base_level"]
    rng = np.random.default_rng(4)
    e = [rng.standard_normal((g.info["nblocks"], n, n, n)) for _ in range(3)]
    f = ot.block_op(4, e, L, scale=0.7)
    div = ot.block_op(2, f, L)
    sl = (slice(None), slice(1, -2), slice(1, -2), slice(1, -2))
    assert np.abs(div[sl]).max() < 1e-9 * np.abs(f[0]).max()


def test_gradient_of_linear_field(uni):
    g, ot = uni
    n, L = g.tree.n, g.info["base_level"]
    dx = g.info["dx_base"]
    phi = _poly_field(g.arrays, n, dx, lambda X, Y, Z: 2 * X - 3 * Y + 5 * Z, L)
    gx, gy, gz = ot.block_op(1, phi, L)
    assert np.allclose(_interior(gx, g.arrays, L), 2 / dx) and np.allclose(_interior(gy, g.arrays, L), -3 / dx)
    assert np.allclose(_interior(gz, g.arrays, L), 5 / dx)


@pytest.mark.parametrize("kind,support", [(0, 2.0), (1, 2.5), (2, 1.5)])
def test_delta_moments(oracle_mod, kind, support):
    """Partition of unity and zero first moment for every shift; support radius.
    yang4 is evaluated as the reference does, whose outer branch multiplies
    asin(sqrt(2)(r-2)) by the integer quotient 1/16 = 0 (ib.hpp:700-701): with
    that term restored the moments hold, which pins both the formula and the
    quirk."""
    d = oracle_mod.delta_1d

    def w_of(p):
        v = d(kind, p)
        r = abs(p)
        if kind == 1 and 1.5 < r <= 2.5:
            v += np.arcsin(np.sqrt(2) * (r - 2)) / 16.0
        return v

    for x in np.linspace(0.0, 1.0, 11):
        pts = np.arange(-4, 6) - x
        w = np.array([w_of(p) for p in pts])
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs((w * pts).sum()) < 1e-14
    assert d(kind, support + 1e-9) == 0.0 and w_of(support - 0.3) > 0.0


def test_roma_closed_form(oracle_mod):
    # roma(0) = 2/3, roma(1) = 1/6 (Roma, Peskin & Berger 1999)
    assert abs(oracle_mod.delta_1d(2, 0.0) - 2.0 / 3.0) < 1e-15
    assert abs(oracle_mod.delta_1d(2, 1.0) - 1.0 / 6.0) < 1e-15


@pytest.fixture(scope="module")
def sphere(oracle_mod):
    s = setups.IbSphere(npoints=200, block_extent=8, nref=1, half_width=0.5)
    ot = oracle_mod.Tree(s.arrays, s.info, s.tree.n)
    ib = oracle_mod.Ib(s.points, s.nref, s.delta, 4.5, s.safety_dis)
    return s, ot, ib


def test_ib_flags_and_masks(sphere):
    s, ot, ib = sphere
    a, L = s.arrays, s.info["depth"] - 1
    fl = a["flags"]
    is_ib, ext = (fl & 16) != 0, (fl & 32) != 0
    assert is_ib.any() and np.all(a["level"][is_ib] == L) and np.all(ext[is_ib])
    assert np.all((fl[is_ib] & 1) != 0)  # is_ib only on leaves
    for idx, sbit, tbit in ((1, is_ib, is_ib | ext), (2, is_ib | ext, is_ib)):
        S = ((a["mask_src"] >> np.uint64(idx)) & np.uint64(1)) == 1
        T = ((a["mask_tgt"] >> np.uint64(idx)) & np.uint64(1)) == 1
        finest = a["level"] == L
        # base-level initialisation (fmm.hpp:70-144); clean_no_inf may only clear sources
        assert np.array_equal(T[finest], tbit[finest])
        assert np.all(~S[finest] | sbit[finest])


def test_host_weights_match_oracle_bitwise(sphere):
    s, ot, ib = sphere
    st = s.tree.ib_stencil()
    n, nb = s.tree.n, s.info["nblocks"]
    rng = np.random.default_rng(1)
    force = rng.standard_normal((s.npoints, 3))
    fv = force.reshape(-1)
    fields = [np.zeros(nb * n ** 3) for _ in range(3)]
    for r in range(len(st["spread_ptr"]) - 1):
        acc = 0.0
        for k in range(st["spread_ptr"][r], st["spread_ptr"][r + 1]):
            acc = acc + fv[st["spread_col"][k]] * st["spread_w"][k]
        c, q = st["spread_dst"][r]
        fields[c][q] = acc
    ref = ib.smear(ot, force)
    for c in range(3):
        assert np.array_equal(fields[c], ref[c].reshape(-1))
    u = [rng.standard_normal(nb * n ** 3) for _ in range(3)]
    out = np.zeros(3 * s.npoints)
    for r in range(3 * s.npoints):
        acc = 0.0
        for k in range(st["interp_ptr"][r], st["interp_ptr"][r + 1]):
            acc = acc + u[r % 3][st["interp_node"][k]] * st["interp_w"][k]
        out[r] = acc
    refp = ib.project(ot, [x.reshape(nb, n, n, n) for x in u]).reshape(-1)
    assert np.array_equal(out, refp)
    assert np.array_equal(st["owned"], ib.owned(ot))


def test_smearing_conserves_force_and_centres_on_face(sphere):
    """Sum of spread weights = 1 and their centroid = X (own axis) or X - 1/2
    (cell-centred axes): the face data location of ib_smearing."""
    s, ot, ib = sphere
    n, L = s.tree.n, s.info["depth"] - 1
    a = s.arrays
    i = 17
    force = np.zeros((s.npoints, 3))
    force[i] = [1.0, 1.0, 1.0]
    spread = ib.smear(ot, force)
    rl = L - s.info["base_level"]
    X = s.points[i] * 2 ** rl / s.dx_base
    idx = np.arange(n) - 1
    for c in range(3):
        tot, mom = 0.0, np.zeros(3)
        for b in np.nonzero(np.any(spread[c] != 0, axis=(1, 2, 3)))[0]:
            bb = a["bbase"][b]
            Z, Y, Xg = np.meshgrid(bb[2] + idx, bb[1] + idx, bb[0] + idx, indexing="ij")
            w = spread[c][b]
            tot += w.sum()
            mom += [(w * Xg).sum(), (w * Y).sum(), (w * Z).sum()]
        assert abs(tot - 1.0) < 1e-13
        want = X - np.array([0.0 if d == c else 0.5 for d in range(3)])
        assert np.allclose(mom / tot, want, atol=1e-12)


def test_smear_project_adjoint(sphere):
    s, ot, ib = sphere
    rng = np.random.default_rng(5)
    f = rng.standard_normal((s.npoints, 3))
    n, nb = s.tree.n, s.info["nblocks"]
    u = [rng.standard_normal((nb, n, n, n)) for _ in range(3)]
    lhs = np.sum(ib.project(ot, u) * f)
    sm = ib.smear(ot, f)
    rhs = sum(np.sum(sm[c] * u[c]) for c in range(3))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
