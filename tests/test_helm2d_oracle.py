"""The oracle's screened 2-D LGF of a spanwise Fourier mode and its direct
convolution (CPU): pinned by the origin value's hypergeometric closed form
(helmholtz_lookup.hpp:114-118), an adaptive Gauss-Kronrod quadrature of the
same integral, the lattice identity (Lap_h - c^2) value = delta, convergence
of the node count, and scipy's convolve2d with the mirrored table."""
import numpy as np
import pytest
from scipy import integrate, signal, special

from oracle import oracle as O

CS = (0.05, 0.3, 1.0, 3.0)


def _g_quad(c, n, m):
    def f(x):
        a = 4.0 + c * c - 2.0 * np.cos(x)
        root = np.sqrt(a * a - 4.0)
        return np.cos(m * x) * ((a + root) / 2.0) ** (-n) / root

    v, _ = integrate.quad(f, 0.0, np.pi, epsabs=1e-15, epsrel=1e-14, limit=2000)
    return v / np.pi


def _mirror(t):
    """Table of offsets 0.. -> full kernel of offsets -(n-1)..(n-1)."""
    return np.block([[t[:0:-1, :0:-1], t[:0:-1, :]], [t[:, :0:-1], t]])


@pytest.mark.parametrize("c", CS)
def test_origin_equals_hypergeometric_closed_form(c):
    z = (2.0 / (2.0 + c * c / 2.0)) ** 2
    origin = 0.5 / (2.0 + c * c / 2.0) * special.hyp2f1(0.5, 0.5, 1.0, z)
    got = -O.helm2d_table(c, 1, 1)[0, 0]
    assert abs(got - origin) <= 1e-13 * origin


@pytest.mark.parametrize("c", CS)
def test_values_match_adaptive_quadrature_and_are_symmetric(c):
    t = O.helm2d_table(c, 12, 9)
    for (i, j) in [(0, 0), (1, 0), (0, 1), (3, 2), (7, 5), (11, 8), (11, 0)]:
        assert abs(-t[j, i] - _g_quad(c, max(i, j), min(i, j))) <= 1e-13 * abs(t[0, 0]), (i, j)
    q = O.helm2d_nodes(c, 21)
    L = O.lib()
    assert L.oracle_helm2d_lgf(c, 3, 7, q) == L.oracle_helm2d_lgf(c, 7, 3, q) == L.oracle_helm2d_lgf(c, -7, 3, q)


@pytest.mark.parametrize("c", CS)
def test_lattice_identity(c):
    # value(n, m) = -G and (-Lap_h + c^2) G = delta, so (Lap_h - c^2) value = delta
    n = 16
    full = _mirror(O.helm2d_table(c, n, n))
    lap = full[1:-1, :-2] + full[1:-1, 2:] + full[:-2, 1:-1] + full[2:, 1:-1] - (4.0 + c * c) * full[1:-1, 1:-1]
    delta = np.zeros_like(lap)
    delta[n - 2, n - 2] = 1.0
    assert np.max(np.abs(lap - delta)) < 1e-12


def test_node_count_converged():
    for c in (1e-2, 0.3):
        q = O.helm2d_nodes(c, 40)
        a, b = O.helm2d_table(c, 20, 20, q), O.helm2d_table(c, 20, 20, 2 * q)
        # what is left is summation rounding: terms near x = 0 reach 1 / (2c)
        assert np.max(np.abs(a - b)) <= 1e-13 * abs(a[0, 0])


def test_convolution_equals_convolve2d_with_mirrored_table():
    rng = np.random.default_rng(11)
    nx, ny, c = 13, 9, 0.4
    t = O.helm2d_table(c, nx, ny)
    src = rng.standard_normal((ny, nx))
    ref = 0.5 * signal.convolve2d(src, _mirror(t), mode="full")[ny - 1:2 * ny - 1, nx - 1:2 * nx - 1]
    got = O.helm2d_convolve(t, src, 0.5)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    # interior cells satisfy (Lap_h - c^2) u = scale * src for the free-space sum
    lap = got[1:-1, :-2] + got[1:-1, 2:] + got[:-2, 1:-1] + got[2:, 1:-1] - (4.0 + c * c) * got[1:-1, 1:-1]
    np.testing.assert_allclose(lap, 0.5 * src[1:-1, 1:-1], atol=1e-12)
