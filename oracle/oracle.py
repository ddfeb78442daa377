"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg;
never by the product package.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None
_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


class OracleIb(ctypes.Structure):
    _fields_ = [("npoints", _I), ("x", _P), ("ib_level", _I), ("delta", _I), ("ddf_radius", _D),
                ("safety_dis", _D)]


class OracleTree(ctypes.Structure):
    _fields_ = [
        ("nb", _I), ("n", _I), ("base_level", _I), ("depth", _I), ("dx_base", _D),
        ("level", _P), ("coord", _P), ("bbase", _P), ("flags", _P), ("parent", _P), ("children", _P),
        ("neighbor", _P), ("influence", _P), ("influence_num", _P), ("mask_src", _P), ("mask_tgt", _P),
    ]


def build():
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_lgf_table.argtypes = [_I, _P]
        L.oracle_helm2d_lgf.argtypes = [_D, _I, _I, _I]
        L.oracle_helm2d_lgf.restype = _D
        L.oracle_helm2d_convolve.argtypes = [_P, _I, _I, _P, _P, _D]
        L.oracle_lgf_get.argtypes = [_P, _I, _I, _I, _I]
        L.oracle_lgf_get.restype = _D
        L.oracle_lgf_asym.argtypes = [_I, _I, _I]
        L.oracle_lgf_asym.restype = _D
        L.oracle_ive.argtypes = [_D, _I, _P]
        L.oracle_lagrange_matrix.argtypes = [_I, _I, _I, _P]
        L.oracle_direct_sum.argtypes = [_P, _I, _I, _P, _P, _I, _P, _D, _P]
        L.oracle_block_conv.argtypes = [_P, _I, _I, _P, _P, _I, _D, _P]
        L.oracle_apply_lgf.argtypes = [ctypes.POINTER(OracleTree), _P, _I, _P, _P, _I, _I]
        L.oracle_apply_if.argtypes = [ctypes.POINTER(OracleTree), _P, _P, _D, _I, _I]
        L.oracle_lgf_baseline.argtypes = [ctypes.POINTER(OracleTree), _P, _I, _P, _P, _I, _I, _I, _P]
        L.oracle_apply_lgf_mesh.argtypes = [ctypes.POINTER(OracleTree), _P, _I, _P, _P, _I, _I]
        L.oracle_apply_if_mesh.argtypes = [ctypes.POINTER(OracleTree), _P, _P, _D, _I, _I]
        L.oracle_mesh_transfer.argtypes = [_I, _I, _I, _I, _I, _P, _P]
        L.oracle_bx_sample.argtypes = [ctypes.POINTER(OracleTree), _P, _I, _P, _I, _I, _I, _P, _I, _P, _P]
        L.oracle_bx_sample.restype = ctypes.c_long
        T, B = ctypes.POINTER(OracleTree), ctypes.POINTER(OracleIb)
        L.oracle_apply_lgf_masked.argtypes = [T, _P, _I, _P, _P, _I, _I, _I]
        L.oracle_apply_if_masked.argtypes = [T, _P, _P, _D, _I, _I, _I]
        L.oracle_block_op.argtypes = [T, _I, _P, _P, _P, _P, _P, _P, _I, _D]
        L.oracle_buffer_exchange.argtypes = [T, _P, _I]
        L.oracle_clean_boundary.argtypes = [T, _P, _I, _I, _I, _I]
        L.oracle_smooth_boundary.argtypes = [T, _P, _I, _I]
        L.oracle_nonlinear.argtypes = [T] + [_P] * 9 + [_I, _D]
        L.oracle_up_and_down.argtypes = [T, _P, _P, _P, _I]
        L.oracle_ifherk_step_ib.argtypes = [T, B, _P, _I, _P, _P, _P, _P, _D, _D, _P, _I, _P, _P, _D, _I]
        L.oracle_ifherk_step_ib_points.argtypes = [T, B, _P, _I, _P, _P, _P, _P, _D, _D, _P, _I, _P, _P, _D, _I]
        L.oracle_ifherk_step.argtypes = [T, _P, _I, _P, _P, _P, _P, _D, _D, _P, _I]
        L.oracle_delta_1d.argtypes = [_I, _D]
        L.oracle_delta_1d.restype = _D
        L.oracle_ib_smear.argtypes = [T, B, _P, _P, _P, _P]
        L.oracle_ib_project.argtypes = [T, B, _P, _P, _P, _P]
        L.oracle_ib_owned.argtypes = [T, B, _P]
        L.oracle_ib_apply.argtypes = [T, B, _P, _I, _P, _P, _D]
        L.oracle_ib_cg.argtypes = [T, B, _P, _I, _P, _P, _D, _D, _I, _P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


_TABLES = {}


def lgf_table(nmax=100):
    if nmax not in _TABLES:
        g = np.zeros((nmax + 1) * (nmax + 2) * (nmax + 3) // 6, np.float64)
        lib().oracle_lgf_table(nmax, _p(g))
        _TABLES[nmax] = g
    return _TABLES[nmax]


def lgf_get(x, y, z, nmax=100):
    return lib().oracle_lgf_get(_p(lgf_table(nmax)), nmax, int(x), int(y), int(z))


def lgf_asym(x, y, z):
    return lib().oracle_lgf_asym(int(x), int(y), int(z))


def ive(x, kmax):
    out = np.zeros(kmax + 1)
    lib().oracle_ive(float(x), int(kmax), _p(out))
    return out


def helm2d_nodes(c, span):
    """Trapezoid intervals the device table uses for screening c and offsets
    n + m <= span: ceil(span + 45 / acosh(1 + c^2/2)) rounded up to 256."""
    q = int(np.ceil(span + 45.0 / np.arccosh(1.0 + 0.5 * c * c)))
    return (q + 255) & ~255


def helm2d_table(c, nx, ny, nodes=None):
    """Table [ny, nx] of the screened 2-D LGF values (-G) of screening c."""
    q = helm2d_nodes(c, nx + ny) if nodes is None else int(nodes)
    L = lib()
    return np.array([[L.oracle_helm2d_lgf(float(c), i, j, q) for i in range(nx)] for j in range(ny)])


def helm2d_convolve(table, src, scale=1.0):
    """Direct convolution of one mode's patch src [ny, nx] with its table."""
    table = np.ascontiguousarray(table, np.float64)
    src = np.ascontiguousarray(src, np.float64)
    ny, nx = src.shape
    dst = np.zeros_like(src)
    lib().oracle_helm2d_convolve(_p(table), nx, ny, _p(src), _p(dst), float(scale))
    return dst


def lagrange_matrix(n, cap, rp):
    A = np.zeros((n, 2 * n))
    lib().oracle_lagrange_matrix(int(n), int(cap), int(rp), _p(A))
    return A


def mesh_transfer(block, mesh, comp, op, child, out=None):
    """One block's level transfer for a field of mesh type `mesh` and field
    index `comp`: op 0 restricts child block `child` into `out` (accumulating),
    op 1 interpolates the parent block into child `child`."""
    b = np.ascontiguousarray(block, np.float64)
    o = np.zeros_like(b) if out is None else out
    assert lib().oracle_mesh_transfer(b.shape[0], int(mesh), int(comp), int(op), int(child), _p(b), _p(o)) == 0
    return o


def remap(old_arrays, new_tree, fields, mesh, comp):
    """The fields after an adaptation, as Ifherk::adapt fills them
    (ifherk.hpp:1047-1074): every block kept by key, the others zero; then per
    field the halo fills of every field at levels depth-2 .. base, and every
    parent of a new leaf (refinement level > 0, not a leaf before; in key order
    of the new leaves, repeats dropped) interpolated into all its children
    (nli_intrp_node, cell_center_nli_intrp.hpp:310-420).  new_tree is an
    oracle Tree of the adapted tree; returns the new field arrays and the
    slots of the parents interpolated."""
    a, info, n = new_tree.a, new_tree.info, new_tree.n
    old_slot = {int(k): s for s, k in enumerate(old_arrays["key"])}
    old_leaves = {int(k) for k, f in zip(old_arrays["key"], old_arrays["flags"]) if f & 1}
    out = []
    for f in fields:
        g = np.zeros((info["nblocks"], n, n, n))
        for s, k in enumerate(a["key"]):
            o = old_slot.get(int(k))
            if o is not None:
                g[s] = f[o]
        out.append(g)
    bl = info["base_level"]
    new = [s for s in range(info["nblocks"]) if (a["flags"][s] & 1) and a["level"][s] > bl
           and int(a["key"][s]) not in old_leaves]
    new.sort(key=lambda s: int(a["key"][s]))
    parents = []
    for s in new:
        pa = int(a["parent"][s])
        if pa >= 0 and (not parents or parents[-1] != pa):
            parents.append(pa)
    for i in range(len(out)):
        for l in range(info["depth"] - 2, bl - 1, -1):
            for g in out:
                new_tree.buffer_exchange(g, l)
        for pa in parents:
            for c in range(8):
                ch = int(a["children"][pa][c])
                if ch >= 0:
                    out[i][ch] = mesh_transfer(out[i][pa], mesh[i], comp[i], 1, c)
    return out, parents


def direct_sum(src_xyz, q, tgt_xyz, scale, nmax=100):
    sx = np.ascontiguousarray(src_xyz, np.int32)
    tx = np.ascontiguousarray(tgt_xyz, np.int32)
    qq = np.ascontiguousarray(q, np.float64)
    phi = np.zeros(len(tx))
    lib().oracle_direct_sum(_p(lgf_table(nmax)), nmax, len(sx), _p(sx), _p(qq), len(tx), _p(tx), float(scale), _p(phi))
    return phi


def block_conv(src, shift, step=1, scale=1.0, nmax=100):
    n = src.shape[0]
    s = np.ascontiguousarray(src, np.float64)
    t = np.zeros_like(s)
    sh = np.ascontiguousarray(shift, np.int32)
    lib().oracle_block_conv(_p(lgf_table(nmax)), nmax, n, _p(s), _p(sh), int(step), float(scale), _p(t))
    return t


class Tree:
    """Oracle view of an exported block table (dict of numpy arrays from iblgf_amd.Tree.export)."""

    def __init__(self, arrays, info, n):
        self.a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
        self.info = info
        self.n = n
        a = self.a
        self.s = OracleTree(
            info["nblocks"], n, info["base_level"], info["depth"], info["dx_base"],
            _p(a["level"]), _p(a["coord"]), _p(a["bbase"]), _p(a["flags"]), _p(a["parent"]), _p(a["children"]),
            _p(a["neighbor"]), _p(a["influence"]), _p(a["influence_num"]), _p(a["mask_src"]), _p(a["mask_tgt"]),
        )

    def apply_lgf(self, src, faithful=False, intrp_order=5, nmax=100):
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        rc = lib().oracle_apply_lgf(ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(s), _p(dst),
                                    1 if faithful else 0, int(intrp_order))
        assert rc == 0
        return dst

    def lgf_baseline(self, src, ranks, nsolves, faithful=False, nmax=100):
        """1 + nsolves apply_lgf solves with the kernel spectra kept; returns
        (seconds of each warm solve, last result).  ranks > 0 splits each pass's
        targets over that many workers by contiguous load (server.hpp:127-170)."""
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        sec = np.zeros(max(int(nsolves), 1))
        rc = lib().oracle_lgf_baseline(ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(s), _p(dst),
                                       1 if faithful else 0, int(ranks), int(nsolves), _p(sec))
        assert rc == 0
        return sec[:int(nsolves)], dst

    def apply_lgf_mesh(self, src, mesh, comp, nmax=100):
        """apply_lgf of component `comp` of a field of mesh type `mesh` (0 cell, 1 face, 2 edge)."""
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        rc = lib().oracle_apply_lgf_mesh(ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(s), _p(dst), int(mesh),
                                         int(comp))
        assert rc == 0
        return dst

    def apply_if_mesh(self, src, alpha_base, mesh, comp):
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        rc = lib().oracle_apply_if_mesh(ctypes.byref(self.s), _p(s), _p(dst), float(alpha_base), int(mesh), int(comp))
        assert rc == 0
        return dst

    def apply_if(self, src, alpha_base, faithful=False, intrp_order=5):
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        rc = lib().oracle_apply_if(ctypes.byref(self.s), _p(s), _p(dst), float(alpha_base), 1 if faithful else 0,
                                   int(intrp_order))
        assert rc == 0
        return dst

    def bx_sample(self, fmm_s, L, nl, targets, faithful=True, nmax=100):
        fs = np.ascontiguousarray(fmm_s, np.float64)
        tg = np.ascontiguousarray(targets, np.int32)
        out = np.zeros((len(tg),) + fs.shape[1:])
        sec = np.zeros(1)
        pairs = lib().oracle_bx_sample(ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(fs), int(L), int(nl),
                                       len(tg), _p(tg), 1 if faithful else 0, _p(out), _p(sec))
        return pairs, float(sec[0]), out

    # -- block operators and immersed boundary --------------------------------
    def apply_lgf_masked(self, src, mask_type, nmax=100):
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        lib().oracle_apply_lgf_masked(ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(s), _p(dst),
                                      int(mask_type), 0, 5)
        return dst

    def apply_if_masked(self, src, alpha_base, mask_type):
        s = np.ascontiguousarray(src, np.float64)
        dst = np.zeros_like(s)
        lib().oracle_apply_if_masked(ctypes.byref(self.s), _p(s), _p(dst), float(alpha_base), int(mask_type), 0, 5)
        return dst

    def block_op(self, op, src, level, scale=1.0, dst=None):
        """src: one field or a list of 3; returns the output field(s) (zeros outside the level's interiors
        unless dst is given)."""
        srcs = [np.ascontiguousarray(x, np.float64) for x in (src if isinstance(src, (list, tuple)) else [src])]
        nd = 1 if op in (0, 2) else 3
        outs = [np.zeros_like(srcs[0]) for _ in range(nd)] if dst is None else dst
        sp = [_p(x) for x in srcs] + [None] * (3 - len(srcs))
        dp = [_p(x) for x in outs] + [None] * (3 - len(outs))
        lib().oracle_block_op(ctypes.byref(self.s), int(op), *sp, *dp, int(level), float(scale))
        return outs[0] if nd == 1 else outs

    def smooth_boundary(self, f, level, leaf_only):
        lib().oracle_smooth_boundary(ctypes.byref(self.s), _p(f), int(level), 1 if leaf_only else 0)

    def nonlinear(self, u, w, level, scale=1.0, dst=None):
        """Operator::nonlinear on the blocks of one level (face u, edge w)."""
        u = [np.ascontiguousarray(x, np.float64) for x in u]
        w = [np.ascontiguousarray(x, np.float64) for x in w]
        d = [np.zeros_like(u[0]) for _ in range(3)] if dst is None else dst
        lib().oracle_nonlinear(ctypes.byref(self.s), *[_p(x) for x in u + w + d], int(level), float(scale))
        return d

    def up_and_down(self, f, mesh=1):
        """Ifherk::up_and_down of three components (returns new arrays)."""
        f = [np.array(x, np.float64, order="C", copy=True) for x in f]
        lib().oracle_up_and_down(ctypes.byref(self.s), *[_p(x) for x in f], int(mesh))
        return f

    def ifherk_step(self, u, dt, Re, frame=None, refresh=0, nmax=100):
        """One IF-HERK time step of the face velocity u (three arrays); returns (u_new, p).
        refresh 1 opens it with pad_velocity and the adaptation-correction step, 2 with pad_velocity only."""
        u = [np.array(x, np.float64, order="C", copy=True) for x in u]
        p = np.zeros_like(u[0])
        fr = None if frame is None else np.ascontiguousarray(np.broadcast_to(np.asarray(frame, np.float64), (3, 3)))
        rc = lib().oracle_ifherk_step(ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, *[_p(x) for x in u], _p(p),
                                      float(dt), float(Re), None if fr is None else _p(fr), int(refresh))
        assert rc == 0
        return u, p

    def buffer_exchange(self, f, level):
        lib().oracle_buffer_exchange(ctypes.byref(self.s), _p(f), int(level))
        return f

    def clean_boundary(self, f, level, kind, leaf_only=True, width=2):
        lib().oracle_clean_boundary(ctypes.byref(self.s), _p(f), int(level), int(kind), 1 if leaf_only else 0,
                                    int(width))
        return f


class Ib:
    """Immersed-boundary points for the oracle (IB::init semantics)."""

    def __init__(self, points, ib_level, delta=0, ddf_radius=4.5, safety_dis=4.0):
        self.x = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        self.s = OracleIb(self.x.shape[0], _p(self.x), int(ib_level), int(delta), float(ddf_radius),
                          float(safety_dis))

    def smear(self, tree, force):
        f = np.ascontiguousarray(force, np.float64).reshape(-1)
        shape = (tree.info["nblocks"], tree.n, tree.n, tree.n)
        out = [np.zeros(shape) for _ in range(3)]
        lib().oracle_ib_smear(ctypes.byref(tree.s), ctypes.byref(self.s), _p(f), *[_p(o) for o in out])
        return out

    def project(self, tree, fields):
        fl = [np.ascontiguousarray(x, np.float64) for x in fields]
        out = np.zeros((self.x.shape[0], 3))
        lib().oracle_ib_project(ctypes.byref(tree.s), ctypes.byref(self.s), *[_p(x) for x in fl], _p(out))
        return out

    def owned(self, tree):
        o = np.zeros(self.x.shape[0], np.int32)
        lib().oracle_ib_owned(ctypes.byref(tree.s), ctypes.byref(self.s), _p(o))
        return o

    def apply(self, tree, fin, alpha, nmax=100):
        f = np.ascontiguousarray(fin, np.float64).reshape(-1)
        out = np.zeros((self.x.shape[0], 3))
        lib().oracle_ib_apply(ctypes.byref(tree.s), ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(f), _p(out),
                              float(alpha))
        return out

    def ifherk_step(self, tree, u, dt, Re, force_prev, frame=None, refresh=0, bc=None, cg_tol=1e-3, cg_maxit=100,
                    nmax=100, bc_points=None):
        """One IF-HERK step with these IB points; force_prev [3 stages, N, 3]
        is updated in place; bc [3 stages, 3] (constant) or bc_points
        [3 stages, N, 3] (per point) is the boundary velocity; returns (u_new, p)."""
        u = [np.array(x, np.float64, order="C", copy=True) for x in u]
        p = np.zeros_like(u[0])
        assert force_prev.dtype == np.float64 and force_prev.flags["C_CONTIGUOUS"]
        assert force_prev.size == 9 * self.x.shape[0]
        fr = None if frame is None else np.ascontiguousarray(np.broadcast_to(np.asarray(frame, np.float64), (3, 3)))
        if bc_points is not None:
            bp = np.ascontiguousarray(bc_points, np.float64).reshape(3, -1, 3)
            assert bp.shape[1] == self.x.shape[0]
            rc = lib().oracle_ifherk_step_ib_points(ctypes.byref(tree.s), ctypes.byref(self.s), _p(lgf_table(nmax)),
                                                    nmax, *[_p(x) for x in u], _p(p), float(dt), float(Re),
                                                    None if fr is None else _p(fr), int(refresh), _p(bp),
                                                    _p(force_prev), float(cg_tol), int(cg_maxit))
            assert rc == 0
            return u, p
        b = None if bc is None else np.ascontiguousarray(np.broadcast_to(np.asarray(bc, np.float64), (3, 3)))
        rc = lib().oracle_ifherk_step_ib(ctypes.byref(tree.s), ctypes.byref(self.s), _p(lgf_table(nmax)), nmax,
                                         *[_p(x) for x in u], _p(p), float(dt), float(Re),
                                         None if fr is None else _p(fr), int(refresh), None if b is None else _p(b),
                                         _p(force_prev), float(cg_tol), int(cg_maxit))
        assert rc == 0
        return u, p

    def cg(self, tree, uc, alpha, tol, maxit, f0=None, nmax=100):
        u = np.ascontiguousarray(uc, np.float64).reshape(-1)
        f = np.zeros_like(u) if f0 is None else np.ascontiguousarray(f0, np.float64).reshape(-1).copy()
        res = np.zeros(1)
        it = lib().oracle_ib_cg(ctypes.byref(tree.s), ctypes.byref(self.s), _p(lgf_table(nmax)), nmax, _p(u), _p(f),
                                float(alpha), float(tol), int(maxit), _p(res))
        return f.reshape(-1, 3), it, float(res[0])


def delta_1d(kind, x):
    return lib().oracle_delta_1d(int(kind), float(x))
