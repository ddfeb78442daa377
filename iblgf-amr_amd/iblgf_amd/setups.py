"""Synthetic seeded workloads of the LGF path.

* ``uniform_gaussians`` -- a single-level uniform mesh (4x4x4 blocks of 8^3
  cells by default) with a mean-free sum of isotropic Gaussians.
* ``vortex_rings`` -- the adaptive vortex-ring octree: 8^3 base blocks of 16^3
  cells, refined where the vorticity magnitude over the block grown by two
  cells exceeds ``source_max * refinement_factor**diff_level``, the rule of the
  reference's vortex-ring Poisson test (tests/poisson/vortexrings.hpp:471-520),
  with Gaussian-core rings.

All draws come from SplitMix64 (Steele, Lea & Flood 2014) with
u = (x >> 11) * 2^-53 and Box-Muller normals, consumed in the documented
order, so the oracle and the device see identical inputs.
"""
import math

import numpy as np

from . import FLAG_LEAF, Tree

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.x = int(seed) & MASK64

    def next(self):
        self.x = (self.x + 0x9E3779B97F4A7C15) & MASK64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self):
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def cell_centres(tree_arrays, slot, n, dx_base, base_level):
    """Physical coordinates of every point of block `slot` (buffers included), [n,n,n,3] (z,y,x order)."""
    rl = max(int(tree_arrays["level"][slot]) - base_level, 0)
    dx = dx_base / (2 ** rl)
    b = tree_arrays["bbase"][slot]
    idx = np.arange(n) - 1
    gx = (b[0] + idx + 0.5) * dx
    gy = (b[1] + idx + 0.5) * dx
    gz = (b[2] + idx + 0.5) * dx
    Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
    return np.stack([X, Y, Z], axis=-1)


def _interior_mask(n):
    m = np.zeros((n, n, n), bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


# ---------------------------------------------------------------------------
# Uniform mesh, sum of Gaussians
# ---------------------------------------------------------------------------
class UniformGaussians:
    def __init__(self, seed=1000, nblocks_side=4, block_extent=8, ngauss=8, sigma=2.0, margin=4.0):
        ext = nblocks_side * block_extent
        half = ext // 2
        self.block_extent = block_extent
        self.ext = ext
        self.dx_base = 1.0 / ext
        self.tree = Tree((-half,) * 3, (ext,) * 3, block_extent, self.dx_base,
                         [(-half, -half, -half, ext, ext, ext)])
        self.tree.init_refine(0)
        self.tree.build_masks()
        self.arrays = self.tree.export()
        self.info = self.tree.info()
        rng = SplitMix64(seed)
        self.gauss = []
        for _ in range(ngauss):
            c = [rng.uniform(margin, ext - margin) for _ in range(3)]
            a = rng.normal()
            self.gauss.append((c, a))
        self.sigma = sigma
        self.half = half

    def source(self):
        """Source field [nb,n,n,n]: interior cells of the leaf blocks, mean-free over the mesh."""
        a = self.arrays
        n = self.tree.n
        nb = self.info["nblocks"]
        f = np.zeros((nb, n, n, n))
        inner = _interior_mask(n)
        leaves = np.nonzero(a["flags"] & FLAG_LEAF)[0]
        idx = np.arange(n) - 1
        for s in leaves:
            b = a["bbase"][s]
            # cell-centre position measured from the box corner, in cells
            px = b[0] + idx + self.half + 0.5
            py = b[1] + idx + self.half + 0.5
            pz = b[2] + idx + self.half + 0.5
            Z, Y, X = np.meshgrid(pz, py, px, indexing="ij")
            v = np.zeros((n, n, n))
            for c, amp in self.gauss:
                r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
                v += amp * np.exp(-r2 / (2.0 * self.sigma ** 2))
            f[s][inner] = v[inner]
        tot = sum(f[s][inner].sum() for s in leaves)
        cnt = len(leaves) * self.block_extent ** 3
        mean = tot / cnt
        for s in leaves:
            f[s][inner] -= mean
        return f


# ---------------------------------------------------------------------------
# Adaptive vortex-ring octree
# ---------------------------------------------------------------------------
class VortexRings:
    """Adaptive tree refined around seeded Gaussian-core vortex rings.

    The bounding box is twice the block region, as in the reference's own
    vortex-ring configurations, because base blocks within two blocks of the
    bounding box are never refined (include/iblgf/domain/octree/tree.hpp:346-358).
    """

    def __init__(self, seed=1001, nrings=16, block_extent=16, base_blocks=8, nref=3, refinement_factor=0.9,
                 centre_range=40.0, radius=(8.0, 24.0), core=(2.0, 4.0), circulation=(0.5, 1.5)):
        e = block_extent
        region = base_blocks * e
        half = region // 2
        self.block_extent = e
        self.dx_base = 1.0 / region
        self.nref = nref
        self.refinement_factor = refinement_factor
        rng = SplitMix64(seed)
        rings = []
        for _ in range(nrings):
            c = np.array([rng.uniform(-centre_range, centre_range) for _ in range(3)])
            R = rng.uniform(*radius)
            a = rng.uniform(*core)
            z = rng.uniform(-1.0, 1.0)
            phi = rng.uniform(0.0, 2.0 * math.pi)
            s = math.sqrt(max(0.0, 1.0 - z * z))
            axis = np.array([s * math.cos(phi), s * math.sin(phi), z])
            gam = rng.uniform(*circulation)
            rings.append((c * self.dx_base, R * self.dx_base, a * self.dx_base, axis, gam))
        self.rings = rings
        self.source_max = max(g / (math.pi * a * a) for (_, _, a, _, g) in rings)
        self.tree = Tree((-region,) * 3, (2 * region,) * 3, e, self.dx_base,
                         [(-half, -half, -half, region, region, region)])
        self.tree.init_refine(nref, self._refine)
        self.tree.build_masks()
        self.arrays = self.tree.export()
        self.info = self.tree.info()

    def vorticity(self, P):
        """omega at points P[...,3] (physical units), returns [...,3]."""
        out = np.zeros(P.shape)
        for c, R, a, axis, gam in self.rings:
            d = P - c
            zax = d @ axis
            perp = d - zax[..., None] * axis
            rp = np.sqrt(np.sum(perp * perp, axis=-1))
            rho2 = (rp - R) ** 2 + zax ** 2
            amp = gam / (math.pi * a * a) * np.exp(-rho2 / (a * a))
            et = np.cross(np.broadcast_to(axis, perp.shape), perp)
            with np.errstate(invalid="ignore", divide="ignore"):
                et = np.where(rp[..., None] > 0, et / rp[..., None], 0.0)
            out += amp[..., None] * et
        return out

    def _refine(self, base, ext, rl, diff):
        dx = self.dx_base / (2 ** rl)
        thr = self.source_max * self.refinement_factor ** diff
        # Rings whose core circle is farther from the grown block than the
        # radius where their peak falls below the threshold cannot refine it.
        lo = (np.array(base) - 2 + 0.5) * dx
        hi = (np.array(base) + ext + 1 + 0.5) * dx
        centre = 0.5 * (lo + hi)
        half_diag = 0.5 * float(np.linalg.norm(hi - lo))
        near = False
        for c, R, a, axis, gam in self.rings:
            amp = gam / (math.pi * a * a)
            if amp <= thr:
                continue
            reach = a * math.sqrt(math.log(amp / thr))
            d = centre - c
            zax = float(d @ axis)
            rp = float(np.linalg.norm(d - zax * axis))
            if math.hypot(rp - R, zax) - half_diag < reach:
                near = True
                break
        if not near:
            return False
        i = np.arange(-2, ext + 2)
        gx = (base[0] + i + 0.5) * dx
        gy = (base[1] + i + 0.5) * dx
        gz = (base[2] + i + 0.5) * dx
        Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
        P = np.stack([X, Y, Z], axis=-1)
        w = self.vorticity(P)
        return bool(np.max(np.sqrt(np.sum(w * w, axis=-1))) > thr)

    def source(self, component):
        """Component `component` of the vorticity on the interior cells of every leaf."""
        a = self.arrays
        n = self.tree.n
        nb = self.info["nblocks"]
        f = np.zeros((nb, n, n, n))
        inner = _interior_mask(n)
        base_level = self.info["base_level"]
        for s in np.nonzero(a["flags"] & FLAG_LEAF)[0]:
            P = cell_centres(a, s, n, self.dx_base, base_level)
            w = self.vorticity(P)[..., component]
            f[s][inner] = w[inner]
        return f


# ---------------------------------------------------------------------------
# Immersed-boundary sphere
# ---------------------------------------------------------------------------
def fibonacci_sphere(npoints, R=0.5):
    """IB::read_points, geometry "sphere" (include/iblgf/domain/ib.hpp:179-200)."""
    lam = (1.0 + math.sqrt(5)) / 2.0
    x = np.arange(npoints) + 0.5
    phi = np.arccos(1.0 - 2 * x / npoints)
    theta = 2 * math.pi * x / lam
    return np.stack([R * np.cos(theta) * np.sin(phi), R * np.sin(theta) * np.sin(phi), R * np.cos(phi)], axis=1)


class IbSphere:
    """Sphere of radius R with N seeded quasi-uniform surface points in a
    (nref+1)-level adaptive mesh refined around it.

    N ~ U{5000..20000} (first draw of SplitMix64(seed)) unless given; the
    point spacing is sqrt(pi / (0.866 N)) as the reference's Fibonacci sphere
    (ib.hpp:185), the finest cell is spacing / ibph, and every point gets a
    jitter U(-jitter, jitter) * spacing per coordinate (draws in point order,
    x, y, z).  The base mesh is a cube of 2*half_width, rounded up to an even
    number (at least 4) of base blocks a side.  Blocks whose grown region holds a point are
    refined (Domain::init_refine, domain.hpp:887-891); the safety distance of
    the extended IB region is 5 / (Re dx_ib) + 1 (ib.hpp:82).  The right-hand
    side is the free stream u = (1, 0, 0) at every point (LinSysSolver::ib_solve
    with a uniform boundary velocity, linsys.hpp:112-129).
    """

    def __init__(self, seed=1004, npoints=None, block_extent=16, nref=2, ibph=1.5, half_width=1.0, R=0.5,
                 jitter=0.1, delta=2, Re=1000.0):
        from . import DELTA_ROMA  # noqa: F401  (delta codes: 0 yang3, 1 yang4, 2 roma)

        rng = SplitMix64(seed)
        if npoints is None:
            npoints = 5000 + min(int(rng.uniform() * 15001), 15000)
        self.npoints = int(npoints)
        self.spacing = math.sqrt(math.pi / (0.86602540378 * self.npoints))
        self.dx_ib = self.spacing / ibph
        self.dx_base = self.dx_ib * 2 ** nref
        e = int(block_extent)
        nb = int(math.ceil(2 * half_width / self.dx_base / e))
        nb = max(nb + nb % 2, 4)  # at least 4 base blocks a side: the influence lists need no wrap-around
        ext = nb * e
        base = -ext // 2
        pts = fibonacci_sphere(self.npoints, R)
        jit = np.array([[rng.uniform(-jitter, jitter) for _ in range(3)] for _ in range(self.npoints)])
        self.points = pts + jit * self.spacing
        self.block_extent = e
        self.nref = int(nref)
        self.delta = int(delta)
        self.safety_dis = 5.0 / (Re * self.dx_ib) + 1.0
        self.tree = Tree([base] * 3, [ext] * 3, e, self.dx_base, [[base, base, base, ext, ext, ext]])
        self.tree.set_ib(self.points, self.nref, self.delta, self.safety_dis)
        self.tree.init_refine(self.nref, None)
        self.tree.build_masks()
        self.info = self.tree.info()
        self.arrays = self.tree.export()
        self.uc = np.tile(np.array([1.0, 0.0, 0.0]), (self.npoints, 1))
