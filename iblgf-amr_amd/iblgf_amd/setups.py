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

from . import FLAG_CORRECTION, FLAG_LEAF, Tree

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
                 centre_range=40.0, radius=(8.0, 24.0), core=(2.0, 4.0), circulation=(0.5, 1.5), box_blocks=None,
                 dx_base=None):
        e = block_extent
        region = base_blocks * e
        half = region // 2
        box = (box_blocks if box_blocks is not None else base_blocks) * e
        self.block_extent = e
        self.dx_base = float(dx_base) if dx_base is not None else 1.0 / region
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
        self.params = np.array([len(rings), self.source_max, refinement_factor, self.dx_base]
                               + [v for (c, R, a, axis, gam) in rings for v in (*c, R, a, *axis, gam)])
        self.tree = Tree((-box,) * 3, (2 * box,) * 3, e, self.dx_base,
                         [(-half, -half, -half, region, region, region)])
        self.tree.init_refine_rings(nref, self.params)
        self.tree.build_masks()
        self.arrays = self.tree.export()
        self.info = self.tree.info()

    @staticmethod
    def _ring(ring, P):
        c, R, a, axis, gam = ring
        d = P - c
        zax = d @ axis
        perp = d - zax[..., None] * axis
        rp = np.sqrt(np.sum(perp * perp, axis=-1))
        rho2 = (rp - R) ** 2 + zax ** 2
        amp = gam / (math.pi * a * a) * np.exp(-rho2 / (a * a))
        et = np.cross(np.broadcast_to(axis, perp.shape), perp)
        with np.errstate(invalid="ignore", divide="ignore"):
            et = np.where(rp[..., None] > 0, et / rp[..., None], 0.0)
        return amp[..., None] * et

    def vorticity(self, P):
        """omega at points P[...,3] (physical units), returns [...,3]."""
        out = np.zeros(P.shape)
        for ring in self.rings:
            out += self._ring(ring, P)
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

    def sources(self, components=(0, 1, 2)):
        """The listed vorticity components on the interior cells of every leaf,
        [len(components), nblocks, n, n, n] (evaluated natively,
        iblgf_vortex_rings_vorticity)."""
        return self.tree.vortex_rings_vorticity(self.params, tuple(components))

    def sources_numpy(self, components=(0, 1, 2)):
        """sources() evaluated with numpy, chunks of leaves at a time (the
        restatement the native evaluation is checked against)."""
        a = self.arrays
        n = self.tree.n
        nb = self.info["nblocks"]
        comps = list(components)
        f = np.zeros((len(comps), nb, n, n, n))
        base_level = self.info["base_level"]
        leaves = np.nonzero(a["flags"] & FLAG_LEAF)[0]
        idx = np.arange(1, n - 1) - 1
        for c0 in range(0, len(leaves), 128):
            chunk = leaves[c0:c0 + 128]
            rl = np.maximum(a["level"][chunk].astype(np.int64) - base_level, 0)
            dx = (self.dx_base / 2.0 ** rl)[:, None]
            b = a["bbase"][chunk].astype(np.float64)
            gx = (b[:, 0:1] + idx + 0.5) * dx
            gy = (b[:, 1:2] + idx + 0.5) * dx
            gz = (b[:, 2:3] + idx + 0.5) * dx
            m = n - 2
            P = np.empty((len(chunk), m, m, m, 3))
            P[..., 0] = gx[:, None, None, :]
            P[..., 1] = gy[:, None, :, None]
            P[..., 2] = gz[:, :, None, None]
            w = self.vorticity(P)
            for i, c in enumerate(comps):
                f[i][chunk, 1:-1, 1:-1, 1:-1] = w[..., c]
        return f

    def source(self, component):
        """Component `component` of the vorticity on the interior cells of every leaf."""
        return self.sources((component,))[0]


# ---------------------------------------------------------------------------
# A forest of independent vortex-ring trees (the batched time-step workload)
# ---------------------------------------------------------------------------
class RingForest:
    """Independent adaptive vortex-ring trees of uneven sizes, stepped together.

    From one SplitMix64 stream (seed 1003 for the 64-tree workload), tree i
    draws its own seed (the i-th 64-bit draw) and then a target block count
    floor(exp(U[ln lo, ln hi])); its ring count scales the 16 rings of the
    4290-block single-tree workload to that target (the actual count is
    recorded).  After the trees, the stream draws one integrating-factor
    alpha = c dt / dx^2 ~ U[0.1, 2] per IF-HERK stage (three): the trees step
    with one time step, so a stage's alpha is shared by all of them.
    Every tree has the single-tree workload's box, base level and dx_base.
    """

    def __init__(self, ntrees=64, seed=1003, blocks=(200, 4000), block_extent=16, box_blocks=8, search_steps=8):
        rng = SplitMix64(seed)
        self.members, self.targets, self.seeds = [], [], []
        for _ in range(ntrees):
            tseed = rng.next()
            target = int(math.floor(math.exp(rng.uniform(math.log(blocks[0]), math.log(blocks[1])))))
            self.seeds.append(tseed)
            self.targets.append(target)
            self.members.append(self._sized(tseed, target, block_extent, box_blocks, search_steps))
        self.alphas = [rng.uniform(0.1, 2.0) for _ in range(3)]
        self.trees = [m.tree for m in self.members]
        self.blocks = [m.info["nblocks"] for m in self.members]
        self.leaves = [m.info["nleaves"] for m in self.members]

    @staticmethod
    def _sized(tseed, target, e, box_blocks, steps):
        """A ring tree near `target` blocks.  The base region (4, 6 or 8 blocks
        a side, centred in the shared box) and the ring count (one per 400
        target blocks, 1 to 8) follow the target; then, from 3 refinement levels
        down to 1 (fewer while even the least refinement overshoots), a
        bisection on the refinement threshold source_max * rf^diff over
        ln(-ln rf).  The closest tree is kept."""
        b = 4 if target < 900 else (6 if target < 2500 else 8)
        nrings = min(8, max(1, int(round(target / 400.0))))

        def build(nref, q):
            return VortexRings(seed=tseed, nrings=nrings, block_extent=e, base_blocks=b, box_blocks=box_blocks,
                               dx_base=1.0 / (box_blocks * e), nref=nref, refinement_factor=math.exp(-q),
                               centre_range=40.0 * b / 8, radius=(8.0 * b / 8, 24.0 * b / 8))

        best = None

        def keep(v):
            nonlocal best
            if best is None or abs(v.info["nblocks"] - target) < abs(best.info["nblocks"] - target):
                best = v

        for nref in (3, 2, 1):
            lo, hi = math.log(1e-5), math.log(4.0)  # ln q: fewer blocks at lo
            v = build(nref, math.exp(lo))
            keep(v)
            if v.info["nblocks"] > 1.3 * target and nref > 1:
                continue
            for _ in range(steps):
                mid = 0.5 * (lo + hi)
                v = build(nref, math.exp(mid))
                keep(v)
                if v.info["nblocks"] < target:
                    lo = mid
                else:
                    hi = mid
            break
        return best


# ---------------------------------------------------------------------------
# Manufactured vortex ring with a closed-form solution
# ---------------------------------------------------------------------------
class AnalyticVortexRing:
    """The reference's Poisson acceptance test (tests/poisson/vortexrings.hpp,
    configs tests/poisson/configs/config_1_lv_{0,1}).

    psi = c1 exp(-c2 / (1 - t^2)) inside the torus t < 1, where
    t = sqrt((r - R)^2 + z^2) / R and r = sqrt(x^2 + y^2) (vortexrings.hpp:132-142);
    the source is its Laplacian, so the unbounded solve of Lap phi = source
    should return psi to the scheme's order.  The Laplacian is written here
    from the chain rule in cylindrical coordinates with s = t^2 and
    f(s) = c1 exp(-c2 / (1 - s)):

        Lap psi = f''(s) 4 s / R^2 + f'(s) (4 / R^2 + 2 (r - R) / (r R^2)),
        f' = -c2 q^2 f,  f'' = (c2^2 q^4 - 2 c2 q^3) f,  q = 1 / (1 - s).

    Domain of the configs: Lx = 2.5 over the bounding box (-96)^3 + 192^3
    base cells, so dx_base = 2.5 / 192; 6^3-cell blocks over the central 96^3
    cells; ring c1 = 1e6, c2 = 15, R = 0.125 at the box centre (the origin).
    Cell centres sit at (i + 0.5) dx_level (vortexrings.hpp:372-412).  nlevels
    refinement levels with the reference's rule: refine where |source| over
    the block grown by two cells exceeds source_max * 0.125^diff_level, with
    source_max = |source| on the ring's core circle (vortexrings.hpp:165-178,
    471-520).  ``scale`` multiplies every block count a side (and divides
    dx_base) for refinement studies at a fixed physical problem.
    """

    C1, C2, R = 1.0e6, 15.0, 0.125

    def __init__(self, nlevels=0, scale=1, block_extent=6, half_box=96, region=96, Lx=2.5, refinement_factor=0.125):
        self.block_extent = int(block_extent)
        hb, reg = int(half_box * scale), int(region * scale)
        self.dx_base = Lx / (2 * hb)
        self.refinement_factor = refinement_factor
        self.source_max = abs(float(self.laplacian(np.array([self.R]), np.array([0.0]), np.array([0.0]))[0]))
        self.tree = Tree((-hb,) * 3, (2 * hb,) * 3, self.block_extent, self.dx_base,
                         [(-reg // 2,) * 3 + (reg,) * 3])
        self.tree.init_refine(int(nlevels), self._refine)
        self.tree.build_masks()
        self.arrays = self.tree.export()
        self.info = self.tree.info()

    @classmethod
    def psi(cls, x, y, z):
        r = np.sqrt(x * x + y * y)
        s = ((r - cls.R) ** 2 + z * z) / (cls.R * cls.R)
        out = np.zeros(np.shape(s))
        m = s < 1.0
        out[m] = cls.C1 * np.exp(-cls.C2 / (1.0 - s[m]))
        return out

    @classmethod
    def laplacian(cls, x, y, z):
        R, c2 = cls.R, cls.C2
        r = np.sqrt(x * x + y * y)
        s = ((r - R) ** 2 + z * z) / (R * R)
        out = np.zeros(np.shape(s))
        m = (s < 1.0) & (r > 0.0)
        sm, rm = s[m], r[m]
        q = 1.0 / (1.0 - sm)
        f = cls.C1 * np.exp(-c2 * q)
        f1 = -c2 * q * q * f
        f2 = (c2 * c2 * q ** 4 - 2.0 * c2 * q ** 3) * f
        out[m] = f2 * 4.0 * sm / (R * R) + f1 * (4.0 / (R * R) + 2.0 * (rm - R) / (rm * R * R))
        return out

    def _refine(self, base, ext, rl, diff):
        dx = self.dx_base / (2 ** rl)
        i = np.arange(-2, ext + 2)
        gx = (base[0] + i + 0.5) * dx
        gy = (base[1] + i + 0.5) * dx
        gz = (base[2] + i + 0.5) * dx
        reach = 2.0 * self.R
        if gx[-1] < -reach or gx[0] > reach or gy[-1] < -reach or gy[0] > reach or gz[-1] < -self.R or gz[0] > self.R:
            return False  # the source vanishes outside the torus |z| < R, r < 2R
        Z, Y, X = np.meshgrid(gz, gy, gx, indexing="ij")
        thr = self.source_max * self.refinement_factor ** diff
        return bool(np.max(np.abs(self.laplacian(X, Y, Z))) > thr)

    def _points(self, s, mesh=0, comp=0):
        """Sample positions of block s for a field of mesh type mesh (0 cell,
        1 face, 2 edge) and field index comp: cell centres, moved half a cell
        down on axis comp (face) or on the other two axes (edge)."""
        P = cell_centres(self.arrays, s, self.tree.n, self.dx_base, self.info["base_level"])
        if mesh:
            rl = max(int(self.arrays["level"][s]) - self.info["base_level"], 0)
            dx = self.dx_base / 2 ** rl
            for ax in range(3):
                if (mesh == 1 and ax == comp) or (mesh == 2 and ax != comp):
                    P[..., ax] -= 0.5 * dx
        return P

    def source(self, mesh=0, comp=0):
        """The source on the interior cells of every leaf (vortexrings.hpp:384-410),
        sampled where a field of mesh type `mesh` and field index `comp` sits."""
        n = self.tree.n
        f = np.zeros((self.info["nblocks"], n, n, n))
        inner = _interior_mask(n)
        for s in np.nonzero(self.arrays["flags"] & FLAG_LEAF)[0]:
            P = self._points(s, mesh, comp)
            f[s][inner] = self.laplacian(P[..., 0], P[..., 1], P[..., 2])[inner]
        return f

    def exact(self):
        """psi at every point of every block."""
        n = self.tree.n
        f = np.zeros((self.info["nblocks"], n, n, n))
        for s in range(self.info["nblocks"]):
            P = self._points(s)
            f[s] = self.psi(P[..., 0], P[..., 1], P[..., 2])
        return f

    def finest_blocks(self):
        """Slots of the non-correction leaves at the finest level: where the
        reference takes its L-infinity error (compute_errors skips correction
        blocks, setups/setup_base.hpp:207-224; finest_level_Linf_error_phi_,
        vortexrings.hpp:245-268)."""
        a = self.arrays
        L = self.info["depth"] - 1
        return np.nonzero((a["level"] == L) & ((a["flags"] & FLAG_LEAF) != 0) & ((a["flags"] & FLAG_CORRECTION) == 0))[0]

    def linf_error(self, phi, mesh=0, comp=0):
        """max |phi - psi| over the interiors of finest_blocks() (psi sampled as source(mesh, comp))."""
        sel = self.finest_blocks()
        ex = np.stack([self.psi(*np.moveaxis(self._points(s, mesh, comp), -1, 0)) for s in sel])
        return float(np.max(np.abs(phi[sel][:, 1:-1, 1:-1, 1:-1] - ex[:, 1:-1, 1:-1, 1:-1])))


# ---------------------------------------------------------------------------
# Immersed-boundary sphere
# ---------------------------------------------------------------------------
def mesh_points(tree_arrays, slot, n, dx_base, base_level, mesh=0, comp=0):
    """Physical positions [n,n,n,3] of block `slot` for a field of mesh type
    `mesh` (0 cell, 1 face, 2 edge) and field index `comp`: cell centres moved
    half a cell down on axis comp (face) or on the two other axes (edge)."""
    P = cell_centres(tree_arrays, slot, n, dx_base, base_level)
    if mesh:
        rl = max(int(tree_arrays["level"][slot]) - base_level, 0)
        dx = dx_base / 2 ** rl
        for ax in range(3):
            if (mesh == 1 and ax == comp) or (mesh == 2 and ax != comp):
                P[..., ax] -= 0.5 * dx
    return P


class GaussianFlow:
    """A face velocity u = curl A with A = sum_k a_k exp(-|x - c_k|^2 / s_k^2) e_k
    (Gaussian vector potentials along unit axes e_k), sampled on every point
    of every block where the face components sit: u = sum_k grad(psi_k) x e_k.
    Compactly supported to round-off when the blobs sit well inside the
    domain; the initial condition of the IF-HERK step workloads."""

    def __init__(self, tree, blobs):
        self.tree = tree
        self.arrays = tree.export()
        self.info = tree.info()
        self.blobs = [(np.asarray(c, float), float(s), np.asarray(e, float) / np.linalg.norm(e), float(a))
                      for c, s, e, a in blobs]

    def _at(self, P):
        u = np.zeros(P.shape)
        for c, s, e, a in self.blobs:
            d = P - c
            psi = a * np.exp(-np.sum(d * d, axis=-1) / (s * s))
            grad = d * (-2.0 / (s * s) * psi)[..., None]
            u += np.cross(grad, e)
        return u

    def velocity(self):
        """[3][nb, n, n, n]: component c sampled at the faces normal to axis c."""
        n, nb = self.tree.n, self.info["nblocks"]
        out = [np.zeros((nb, n, n, n)) for _ in range(3)]
        for s in range(nb):
            for c in range(3):
                P = mesh_points(self.arrays, s, n, self.info["dx_base"], self.info["base_level"], 1, c)
                out[c][s] = self._at(P)[..., c]
        return out


class NsAmrLgfRing:
    """The reference's IF-HERK known-answer test (tests/ns_amr_lgf,
    configs/configFile_0 and configFile_1): a fat vortex ring
    (vortex_ring_vor_fat_ic, ns_amr_lgf.hpp:596-619: w_theta = alpha / R^2
    exp(-4 s^2 / (R^2 - s^2)), alpha = 0.54857674, R = 0.5, s the distance to
    the core circle) in the box [-24, 24)^3 base cells with 6^3-cell blocks;
    config 0: dx_base 0.25, no refinement, 4 base steps; config 1: dx_base
    0.5, one refinement level where the vorticity over the block grown by two
    cells exceeds w_max 0.125^diff (refinement, ns_amr_lgf.hpp:757-818), 2
    base steps of 2 finest steps.  Re 1000, cfl 0.35: dt_base = dx_base cfl
    (ifherk.hpp:125), dt = dt_base / 2^nref, so both run 4 steps of 0.0875 to
    T = 0.35.  The initial velocity is -curl^T of the LGF solve of the
    vorticity (ns_amr_lgf.hpp:505-560); after the steps the reference cleans
    the base level (clean_leaf_correction_boundary(base, true)) and compares
    u_0 on the non-correction leaves of the finest level with ref.hdf5
    (cells [-12, 11]^3 at dx 0.25, zero outside) to L_inf EXP_LInf (1e-2 for
    config 0, 3e-4 for config 1)."""

    ALPHA, R = 0.54857674, 0.5
    DT, RE, STEPS = 0.0875, 1000.0, 4

    def __init__(self, config=0):
        self.config = int(config)
        self.dx_base = 0.25 if self.config == 0 else 0.5
        self.EXP_LINF = 1e-2 if self.config == 0 else 3e-4
        self.tree = Tree((-24,) * 3, (48,) * 3, 6, self.dx_base, [(-24, -24, -24, 48, 48, 48)])
        if self.config == 0:
            self.tree.init_refine(0)
        else:
            w_max = abs(self.ALPHA / (self.R * self.R))

            def refine(b, ext, rl, diff):
                dxl = self.dx_base / 2 ** rl
                i = np.arange(b[0] - 2, b[0] + ext + 2)
                j = np.arange(b[1] - 2, b[1] + ext + 2)
                k = np.arange(b[2] - 2, b[2] + ext + 2)
                K, J, I = np.meshgrid(k, j, i, indexing="ij")
                thr = w_max * 0.125 ** diff
                for c in range(3):
                    x = (I + (0.5 if c == 0 else 0.0)) * dxl
                    y = (J + (0.5 if c == 1 else 0.0)) * dxl
                    z = (K + (0.5 if c == 2 else 0.0)) * dxl
                    if np.any(np.abs(self.vorticity_at(x, y, z, c)) > thr):
                        return True
                return False

            self.tree.init_refine(1, refine)
        self.tree.build_masks()
        self.arrays = self.tree.export()
        self.info = self.tree.info()

    @classmethod
    def vorticity_at(cls, x, y, z, comp):
        R2 = cls.R * cls.R
        r = np.sqrt(x * x + y * y)
        s2 = z * z + (r - cls.R) ** 2
        theta = np.arctan2(y, x)
        out = np.zeros(np.shape(x))
        m = s2 < R2
        w = cls.ALPHA / R2 * np.exp(-4.0 * s2[m] / (R2 - s2[m]))
        if comp == 0:
            out[m] = -w * np.sin(theta[m])
        elif comp == 1:
            out[m] = w * np.cos(theta[m])
        return out

    def vorticity(self):
        """[3][nb, n, n, n]: the edge components on every block's interior."""
        n, nb = self.tree.n, self.info["nblocks"]
        inner = _interior_mask(n)
        out = [np.zeros((nb, n, n, n)) for _ in range(3)]
        for s in range(nb):
            for c in range(3):
                P = mesh_points(self.arrays, s, n, self.dx_base, self.info["base_level"], 2, c)
                out[c][s][inner] = self.vorticity_at(P[..., 0], P[..., 1], P[..., 2], c)[inner]
        return out

    def reference_error(self, u0, ref):
        """L_inf of u_0 minus the reference's u_0 (ref dict from
        tests/golden/ns_amr_lgf_ref.npz) over the interiors of the
        non-correction leaves of the finest level."""
        a, n = self.arrays, self.tree.n
        lo, m = int(ref["lo"]), ref["u"].shape[-1]
        L = self.info["depth"] - 1
        err = 0.0
        for s in range(self.info["nblocks"]):
            if a["level"][s] != L or not (a["flags"][s] & FLAG_LEAF) or (a["flags"][s] & FLAG_CORRECTION):
                continue
            b = a["bbase"][s]
            idx = [np.arange(b[d], b[d] + n - 2) - lo for d in range(3)]
            refblk = np.zeros((n - 2,) * 3)
            ok = [(i >= 0) & (i < m) for i in idx]
            sub = ref["u"][0][np.ix_(idx[2][ok[2]], idx[1][ok[1]], idx[0][ok[0]])]
            refblk[np.ix_(ok[2], ok[1], ok[0])] = sub
            err = max(err, float(np.max(np.abs(u0[s][1:-1, 1:-1, 1:-1] - refblk))))
        return err


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


def adapt_level_changes(arrays, base_level, maxnorm, source_max, nref, refinement_factor=0.125,
                        base_threshold=1e-4, hard_max_level=None, ib_add_level=0):
    """The adaptation criterion of the reference's IF-HERK vortex-ring setup
    (NS_AMR_LGF::adapt_level_change and adapt_levle_change_for_field,
    tests/ns_amr_lgf/ns_amr_lgf.hpp:402-495), from per-block maximum norms.

    arrays: the tree's export; maxnorm: dict of per-slot Operator::maxnorm
    arrays for "cell_aux", "correction_tmp" and "edge_aux" (e.g. from
    PoissonSolver.block_norms); source_max: (cell_aux, edge_aux) maxima
    (Ifherk::update_source_max).  Returns {key: level change} for
    Tree.adapt: a leaf's own change, or for a correction block whose field
    asks for refinement, +1 on its parent."""
    hard_max = nref if hard_max_level is None else hard_max_level
    deletion_factor = 0.7

    def field_change(slot, fmax, smax, use_base):
        rl = int(arrays["level"][slot]) - base_level
        flags = int(arrays["flags"][slot])
        is_ib = bool(flags & 16)
        if is_ib and (flags & FLAG_LEAF):
            return 1 if rl < nref + ib_add_level else 0
        smax *= 1.05
        if fmax < 1e-10:
            return -1
        l_aim = int(math.ceil(nref - math.log(fmax / smax) / math.log(refinement_factor)))
        l_del = int(math.ceil(nref - (math.log(fmax / smax) - math.log(deletion_factor)) / math.log(refinement_factor)))
        l_aim = min(l_aim, hard_max)
        if rl == 0 and use_base:
            if fmax > smax * base_threshold:
                l_aim = max(l_aim, 0)
            if fmax > smax * base_threshold * deletion_factor:
                l_del = max(l_del, 0)
        if is_ib:
            return 0
        if l_aim - rl > 0:
            return 1
        if l_del - rl < 0:
            return -1
        return 0

    out = {}
    for s in range(len(arrays["key"])):
        flags = int(arrays["flags"][s])
        leaf, corr = bool(flags & FLAG_LEAF), bool(flags & FLAG_CORRECTION)
        if int(arrays["level"][s]) < base_level or (not leaf and not corr) or (leaf and corr):
            continue
        l1 = l2 = l3 = -1
        if leaf:
            l1 = field_change(s, maxnorm["cell_aux"][s], source_max[0], True)
            l3 = field_change(s, maxnorm["edge_aux"][s], source_max[1], False)
        else:
            l2 = field_change(s, maxnorm["correction_tmp"][s], source_max[0], False)
        l = max(l1, l2, l3)
        if l != 0:
            if leaf:
                out[int(arrays["key"][s])] = l
            elif l2 > 0:
                out[int(arrays["key"][arrays["parent"][s]])] = l2
    return out
