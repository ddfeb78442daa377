"""Screened 2-D LGF tables and their free-space convolution on the device
(needs an MI355X), through the C ABI, against the FP64 oracle: tables within
1e-13 of G(0, 0), convolutions within relative L2 1e-12; at a 256 x 256 patch
the identity (Lap_h - c^2) u = scale * src at interior cells, relative 1e-9."""
import numpy as np
import pytest

from iblgf_amd import IblgfError, helm2d_convolve, helm2d_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _lap(u, c):
    return u[..., 1:-1, :-2] + u[..., 1:-1, 2:] + u[..., :-2, 1:-1] + u[..., 2:, 1:-1] - (4.0 + c * c) * u[..., 1:-1, 1:-1]


def test_tables_match_oracle():
    print("case: tables of 3 modes (c = 0.02, 0.4, 2.5) on a 48 x 40 patch", flush=True)
    cs = np.array([0.02, 0.4, 2.5])
    nx, ny = 48, 40
    t = torch.empty((3, ny, nx), dtype=torch.float64, device="cuda")
    helm2d_table(cs, nx, ny, t)
    got = t.cpu().numpy()
    for k, c in enumerate(cs):
        ref = O.helm2d_table(c, nx, ny)
        assert np.max(np.abs(got[k] - ref)) <= 1e-13 * abs(ref[0, 0]), c


def test_convolution_matches_oracle_per_mode():
    print("case: convolution of 4 modes on a 24 x 17 patch, scale 0.7", flush=True)
    cs = np.array([0.05, 0.3, 1.0, 3.0])
    nx, ny = 24, 17
    rng = np.random.default_rng(21)
    src = rng.standard_normal((4, ny, nx))
    t = torch.empty((4, ny, nx), dtype=torch.float64, device="cuda")
    helm2d_table(cs, nx, ny, t)
    s = torch.from_numpy(src).cuda()
    d = torch.empty_like(s)
    helm2d_convolve(t, s, d, 0.7)
    got, tab = d.cpu().numpy(), t.cpu().numpy()
    for k in range(4):
        ref = O.helm2d_convolve(tab[k], src[k], 0.7)
        assert np.linalg.norm(got[k] - ref) <= 1e-12 * np.linalg.norm(ref), k


@pytest.mark.parametrize("nx,ny", [(64, 3), (72, 5), (300, 2), (1024, 1), (1025, 1)])
def test_row_staged_widths_match_oracle(nx, ny):
    print(f"case: convolution of 2 modes on a {nx} x {ny} patch (row-staged widths and the edges)", flush=True)
    cs = np.array([0.08, 1.5])
    rng = np.random.default_rng(nx + ny)
    src = rng.standard_normal((2, ny, nx))
    t = torch.empty((2, ny, nx), dtype=torch.float64, device="cuda")
    helm2d_table(cs, nx, ny, t)
    s = torch.from_numpy(src).cuda()
    d = torch.empty_like(s)
    helm2d_convolve(t, s, d, -1.25)
    got, tab = d.cpu().numpy(), t.cpu().numpy()
    for k in range(2):
        ref = O.helm2d_convolve(tab[k], src[k], -1.25)
        assert np.linalg.norm(got[k] - ref) <= 1e-12 * np.linalg.norm(ref), k


def test_full_patch_satisfies_screened_poisson():
    print("case: 8 modes on a 256 x 256 patch, (Lap_h - c^2) u = src", flush=True)
    nx = ny = 256
    cs = 2.0 * np.pi * np.arange(1, 9) / 64.0 * (1.0 / 8.0)  # k dx for spanwise modes 1..8
    torch.manual_seed(5)
    s = torch.randn((8, ny, nx), dtype=torch.float64, device="cuda")
    t = torch.empty_like(s)
    d = torch.empty_like(s)
    helm2d_table(cs, nx, ny, t)
    helm2d_convolve(t, s, d, 1.0)
    u, f = d.cpu().numpy(), s.cpu().numpy()
    for k, c in enumerate(cs):
        r = _lap(u[k], c) - f[k, 1:-1, 1:-1]
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(f[k]), (k, c)


def test_argument_errors():
    print("case: argument errors of the 2-D Helmholtz entry points", flush=True)
    t = torch.empty((1, 8, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(IblgfError):
        helm2d_table([0.0], 8, 8, t)  # no screening: the Laplace kernel, not this one
    with pytest.raises(IblgfError):
        helm2d_table([1e-5], 8, 8, t)
    with pytest.raises(IblgfError):
        helm2d_table(np.full(257, 0.5), 1, 1, torch.empty(257, dtype=torch.float64, device="cuda"))
    big = torch.empty((1, 1, 65537), dtype=torch.float64, device="cuda")
    with pytest.raises(IblgfError):
        helm2d_table([0.5], 65537, 1, big)
    with pytest.raises(IblgfError):
        helm2d_convolve(t, t, t)  # source and target must differ
    with pytest.raises(ValueError):
        helm2d_table([0.5], 8, 8, t.float())
