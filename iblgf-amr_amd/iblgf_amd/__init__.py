"""Python binding of libiblgf_gpu.so, the MI355X LGF fast-multipole path.

The classes mirror the reference's interfaces for this path:

* :class:`Tree` -- ``Domain`` / ``Tree`` construction, ``init_refine``,
  ``FmmMaskBuilder`` (include/iblgf/domain/domain.hpp:106-211, 878-943;
  include/iblgf/fmm/fmm.hpp:392-680);
* :class:`PoissonSolver` -- ``PoissonSolver::apply_lgf`` / ``apply_lgf_IF``
  (include/iblgf/solver/poisson/poisson.hpp:130-135, 427-494, 587-894) on one
  GPU with the tree resident, and the block operators of ``domain::Operator``
  (include/iblgf/operators/operators.hpp);
* :class:`IbSystem` -- ``LinSysSolver`` smearing / projection, ``ET_H_S_E``
  and ``CG_solve`` (include/iblgf/solver/linsys/linsys.hpp:170-531).

Everything calls the C ABI declared in ``include/iblgf_gpu.h``.  There is no
CPU fallback: constructing a :class:`PoissonSolver` without the built library
or without a GPU raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiblgf_gpu.so")

# status codes of the C ABI (enum iblgf_status)
IBLGF_OK, IBLGF_ERR_ARG, IBLGF_ERR_STATE, IBLGF_ERR_DEVICE, IBLGF_ERR_UNSUPPORTED, IBLGF_ERR_INTERNAL = range(6)

KERNEL_LAPLACE = 0
KERNEL_IF = 1

PREC_FP64 = 0
PREC_MIXED = 1
PREC_BF16X3 = 2
PREC_FP16X3 = 3
PREC_FP16 = 4
PREC_INT8X4 = 5
PREC_FP16X3_FORWARD = 6
MESH_CELL, MESH_FACE, MESH_EDGE, MESH_VERTEX = range(4)

FLAG_LEAF = 1
FLAG_CORRECTION = 2
FLAG_LEAF_BOUNDARY = 4
FLAG_PHYSICAL = 8
FLAG_IB = 16
FLAG_EXTENDED_IB = 32

OP_LAPLACE, OP_GRADIENT, OP_DIVERGENCE, OP_CURL, OP_CURL_TRANSPOSE = range(5)
CLEAN_LEAF_CORRECTION = 0
CLEAN_IB_REGION = 1
MASK_IB2xIB = 2
MASK_xIB2IB = 3
# the reference's MASK_TYPE (include/iblgf/domain/octree/octant.hpp:75-85)
(FMM_STREAM, FMM_AMR2AMR, FMM_IB2xIB, FMM_xIB2IB, FMM_IB2AMR, FMM_LAPLACIAN_BC, FMM_REF_FOURIER_STREAM,
 FMM_REF_FOURIER_LGF) = range(8)
DELTA_YANG3, DELTA_YANG4, DELTA_ROMA = range(3)

_lib = None

REFINE_FN = ctypes.CFUNCTYPE(
    ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, ctypes.c_int
)

# Every symbol of include/iblgf_gpu.h with (restype, argtypes).
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_U64 = ctypes.c_uint64
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
SYMBOLS = {
    "iblgf_tree_create": (_I, [_IP, _IP, _I, _D, _I, _IP, ctypes.POINTER(_P)]),
    "iblgf_tree_import": (_I, [_IP, _I, _I, _D, _I, ctypes.POINTER(_U64), ctypes.POINTER(ctypes.c_uint32),
                               ctypes.POINTER(_P)]),
    "iblgf_tree_init_refine": (_I, [_P, _I, _I, REFINE_FN, _P]),
    "iblgf_vortex_rings_refine": (_I, [_P, _IP, _I, _I, _I]),
    "iblgf_vortex_rings_vorticity": (_I, [_P, _DP, _I, _DP, _DP, _DP]),
    "iblgf_tree_build_masks": (_I, [_P, _I]),
    "iblgf_tree_adapt": (_I, [_P, _I, ctypes.POINTER(_U64), _IP, _IP]),
    "iblgf_tree_info": (_I, [_P, _IP, _IP, _IP, _IP, _DP, _IP]),
    "iblgf_tree_export": (_I, [_P] + [_P] * 12),
    "iblgf_tree_destroy": (None, [_P]),
    "iblgf_key_make": (_U64, [_I, _I, _I, _I]),
    "iblgf_key_coord": (_I, [_U64, _IP]),
    "iblgf_key_parent": (_U64, [_U64]),
    "iblgf_key_child": (_U64, [_U64, _I]),
    "iblgf_key_influence": (_I, [_U64, ctypes.POINTER(_U64)]),
    "iblgf_lgf_table": (_I, [_I, _DP]),
    "iblgf_lgf_get": (_D, [_I, _I, _I]),
    "iblgf_if_kernel": (_D, [_I, _I, _I, _D]),
    "iblgf_lagrange_matrix": (_I, [_I, _I, _I, _DP]),
    "iblgf_gpu_create": (_I, [_I, ctypes.POINTER(_P)]),
    "iblgf_gpu_destroy": (_I, [_P]),
    "iblgf_gpu_last_error": (ctypes.c_char_p, [_P]),
    "iblgf_last_error": (ctypes.c_char_p, []),
    "iblgf_gpu_tree_upload": (_I, [_P, _P]),
    "iblgf_gpu_field_create": (_I, [_P, _IP]),
    "iblgf_gpu_field_device_ptr": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "iblgf_gpu_field_upload": (_I, [_P, _I, _DP]),
    "iblgf_gpu_field_download": (_I, [_P, _I, _DP]),
    "iblgf_gpu_apply_lgf": (_I, [_P, _I, _I, _I, _D]),
    "iblgf_gpu_apply_lgf_batch": (_I, [_P, _IP, _IP, _I, _I, _D]),
    "iblgf_gpu_set_graphs": (_I, [_P, _I]),
    "iblgf_gpu_set_concurrency": (_I, [_P, _I]),
    "iblgf_gpu_prepare": (_I, [_P, _I, _D]),
    "iblgf_gpu_set_precision": (_I, [_P, _I]),
    "iblgf_gpu_sync": (_I, [_P]),
    "iblgf_gpu_stream": (_I, [_P, ctypes.POINTER(_P)]),
    "iblgf_gpu_counts": (_I, [_P, ctypes.POINTER(ctypes.c_long)]),
    "iblgf_gpu_mac_counts": (_I, [_P, ctypes.POINTER(ctypes.c_long)]),
    "iblgf_gpu_timing_enable": (_I, [_P, _I]),
    "iblgf_gpu_timing_read": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
    "iblgf_gpu_timing_reset": (_I, [_P]),
    "iblgf_tree_set_ib": (_I, [_P, _DP, _I, _I, _I, _D]),
    "iblgf_tree_ib_stencil": (_I, [_P, ctypes.POINTER(ctypes.c_long)] + [_P] * 8),
    "iblgf_gpu_block_op": (_I, [_P, _I, _IP, _IP, _I, _D]),
    "iblgf_gpu_buffer_exchange": (_I, [_P, _I, _I]),
    "iblgf_gpu_level_divergence": (_I, [_P, _IP, _I, _I]),
    "iblgf_gpu_level_gradient": (_I, [_P, _I, _IP, _I]),
    "iblgf_gpu_clean_boundary": (_I, [_P, _I, _I, _I, _I, _I]),
    "iblgf_gpu_smooth_boundary": (_I, [_P, _I, _I, _I]),
    "iblgf_gpu_up_and_down": (_I, [_P, _IP, _I, _I]),
    "iblgf_gpu_nonlinear": (_I, [_P, _IP, _IP, _IP, _I, _D]),
    "iblgf_gpu_ifherk_step": (_I, [_P, _IP, _I, _D, _D, _DP, _I]),
    "iblgf_gpu_ifherk_step_ib": (_I, [_P, _IP, _I, _D, _D, _DP, _I, _DP, _D, _I, _IP]),
    "iblgf_gpu_ib_stage_forces": (_I, [_P, _DP, _I]),
    "iblgf_gpu_apply_lgf_masked": (_I, [_P, _I, _I, _I, _D, _I]),
    "iblgf_gpu_apply_lgf_typed": (_I, [_P, _I, _I, _I, _D, _I]),
    "iblgf_gpu_forest_upload": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "iblgf_gpu_tree_readapt": (_I, [_P, _P, _I, _IP, _IP, _IP, _IP]),
    "iblgf_gpu_coarsen": (_I, [_P, _IP, _I, _I]),
    "iblgf_gpu_field_create_f32": (_I, [_P, _IP]),
    "iblgf_gpu_field_upload_f32": (_I, [_P, _I, ctypes.POINTER(ctypes.c_float)]),
    "iblgf_gpu_field_download_f32": (_I, [_P, _I, ctypes.POINTER(ctypes.c_float)]),
    "iblgf_gpu_solve_sequence": (_I, [_P, _I, _P]),
    "iblgf_gpu_apply_lgf_mesh": (_I, [_P, _IP, _IP, _I, _I, _D, _I]),
    "iblgf_gpu_autotune": (_I, [_P, _IP, _IP, _I, _I, _D, _I, ctypes.c_char_p, _I]),
    "iblgf_gpu_launch_shapes": (_I, [_P, _IP, _IP, _IP]),
    "iblgf_gpu_set_launch_shapes": (_I, [_P, _I, _I]),
    "iblgf_gpu_forest_slots": (_I, [_P, _I, _IP]),
    "iblgf_gpu_apply_lgf_noalloc": (_I, [_P, _I, _I, _I, _D]),
    "iblgf_gpu_helm2d_table": (_I, [_DP, _I, _I, _I, _P, _P]),
    "iblgf_gpu_helm2d_convolve": (_I, [_P, _I, _I, _I, _P, _P, _D, _P]),
    "iblgf_gpu_ib_create": (_I, [_P, _P]),
    "iblgf_gpu_ib_info": (_I, [_P, _IP, _IP, ctypes.POINTER(ctypes.c_long)]),
    "iblgf_gpu_ib_smear": (_I, [_P, _DP, _IP]),
    "iblgf_gpu_ib_project": (_I, [_P, _IP, _DP]),
    "iblgf_gpu_ib_apply": (_I, [_P, _DP, _DP, _D]),
    "iblgf_gpu_ib_cg": (_I, [_P, _DP, _DP, _D, _D, _I, _IP, _DP, _DP]),
    "iblgf_gpu_ib_stats": (_I, [_P, _DP, _IP, _IP, _IP]),
    "iblgf_gpu_ib_precond_create": (_I, [_P, _P, _I, _I, _U64]),
    "iblgf_gpu_ib_precond_collect": (_I, [_P, _I]),
    "iblgf_gpu_ib_precond_train": (_I, [_P, _I, _I, _D, _DP, _IP]),
    "iblgf_gpu_ib_precond_info": (_I, [_P, _IP, _IP, _IP, ctypes.POINTER(ctypes.c_long), _DP]),
    "iblgf_gpu_ib_precond_params": (_I, [_P, _P, ctypes.c_long]),
    "iblgf_gpu_ib_precond_set_params": (_I, [_P, _P, ctypes.c_long]),
    "iblgf_gpu_ib_precond_features": (_I, [_P, _DP, _P]),
    "iblgf_gpu_ib_precond_grad": (_I, [_P, _P, _P, _I, _P, _DP]),
    "iblgf_gpu_ib_fcg": (_I, [_P, _DP, _DP, _D, _D, _I, _I, _IP, _DP, _DP]),
    "iblgf_gpu_ib_switch_iteration": (_I, [_P, _IP]),
    "iblgf_gpu_ib_recycle": (_I, [_P, _I]),
    "iblgf_tree_write_restart": (_I, [_P, ctypes.c_char_p]),
    "iblgf_tree_read_restart": (_I, [ctypes.c_char_p, _IP, _IP, _I, _D, ctypes.POINTER(_P)]),
    "iblgf_chombo_write": (_I, [_P, ctypes.c_char_p, _I, ctypes.POINTER(ctypes.c_char_p), _IP,
                                ctypes.POINTER(_DP), _I, _I, _D, _D]),
    "iblgf_chombo_read": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I, _DP]),
    "iblgf_gpu_chombo_write": (_I, [_P, ctypes.c_char_p, _I, ctypes.POINTER(ctypes.c_char_p), _IP, _IP, _I, _I,
                                    _D, _D]),
    "iblgf_gpu_chombo_read": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p, _I, _I, _IP]),
    "iblgf_gpu_block_norms": (_I, [_P, _IP, _I, _DP, _DP]),
    "iblgf_gpu_ib_set_boundary_velocity": (_I, [_P, _DP]),
}

KERNEL_CLASSES = ("forward_dft", "spectral_mac", "inverse_dft", "transfer", "other")


class IblgfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[iblgf status {code}] {msg}")
        self.code = code


def lib():
    """Load libiblgf_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run iblgf-amr_amd/build.py (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SYMBOLS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, ctx=None):
    if rc != 0:
        L = lib()
        msg = (L.iblgf_gpu_last_error(ctx) if ctx else L.iblgf_last_error()) or b""
        raise IblgfError(rc, msg.decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _ip(v):
    a = (ctypes.c_int * len(v))(*[int(x) for x in v])
    return a


# ---------------------------------------------------------------------------
# Octree
# ---------------------------------------------------------------------------
class Tree:
    """Block octree with refinement, lists and FMM masks (host, C++)."""

    def __init__(self, bd_base, bd_extent, block_extent, dx_base, regions):
        L = lib()
        regs = np.ascontiguousarray(np.asarray(regions, dtype=np.int32).reshape(-1, 6))
        h = ctypes.c_void_p()
        _check(
            L.iblgf_tree_create(
                _ip(bd_base), _ip(bd_extent), int(block_extent), float(dx_base), regs.shape[0],
                regs.ctypes.data_as(_IP), ctypes.byref(h),
            )
        )
        self._h = h
        self.block_extent = int(block_extent)
        self.n = self.block_extent + 2
        self.dx_base = float(dx_base)

    @classmethod
    def from_keys(cls, base_shift, block_extent, base_level, dx_base, keys, flags):
        """A tree handed over by its keys and leaf/correction flags (the reference's own tree)."""
        obj = cls.__new__(cls)
        k = np.ascontiguousarray(keys, dtype=np.uint64)
        f = np.ascontiguousarray(flags, dtype=np.uint32)
        h = ctypes.c_void_p()
        _check(lib().iblgf_tree_import(_ip(base_shift), int(block_extent), int(base_level), float(dx_base), len(k),
                                       k.ctypes.data_as(ctypes.POINTER(_U64)),
                                       f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(h)))
        obj._h = h
        obj.block_extent = int(block_extent)
        obj.n = obj.block_extent + 2
        obj.dx_base = float(dx_base)
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.iblgf_tree_destroy(h)
            self._h = None

    def set_ib(self, points, ib_level, delta=DELTA_YANG3, safety_dis=4.0):
        """Immersed-boundary points (N, 3) at refinement level ib_level (IB::init)."""
        x = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        _check(lib().iblgf_tree_set_ib(self._h, x.ctypes.data_as(_DP), x.shape[0], int(ib_level), int(delta),
                                       float(safety_dis)))
        self.ib_points = x.copy()
        return self

    def ib_stencil(self):
        """Host-built delta weights (the device's): dict of numpy arrays."""
        sz = (ctypes.c_long * 4)()
        _check(lib().iblgf_tree_ib_stencil(self._h, sz, *([None] * 8)))
        ni, nr, ns, npnt = sz[0], sz[1], sz[2], sz[3]
        d = dict(interp_ptr=np.zeros(3 * npnt + 1, np.int32), interp_node=np.zeros(ni, np.int32),
                 interp_w=np.zeros(ni), spread_ptr=np.zeros(nr + 1, np.int32), spread_dst=np.zeros((nr, 2), np.int32),
                 spread_col=np.zeros(ns, np.int32), spread_w=np.zeros(ns), owned=np.zeros(npnt, np.int32))
        _check(lib().iblgf_tree_ib_stencil(self._h, sz, *[_ptr(d[k]) for k in (
            "interp_ptr", "interp_node", "interp_w", "spread_ptr", "spread_dst", "spread_col", "spread_w", "owned")]))
        return d

    def init_refine(self, nref, cond=None, level_up_max=0):
        """Domain::init_refine.  cond(block_base(3,), extent, refinement_level, diff_level) -> bool
        (or 2 to refine without the 2:1 check).  Blocks near immersed-boundary points set with
        set_ib() are refined first, as the reference does."""
        err = []

        def _cb(user, base, ext, rl, diff):
            try:
                if cond is None:
                    return 0
                r = cond((base[0], base[1], base[2]), ext, rl, diff)
                return 2 if (r is not True and r == 2) else (1 if r else 0)
            except Exception as e:  # noqa: BLE001 - surfaced after the call
                err.append(e)
                return 0

        cb = REFINE_FN(_cb)
        _check(lib().iblgf_tree_init_refine(self._h, int(nref), int(level_up_max), cb, None))
        if err:
            raise err[0]
        return self

    def init_refine_rings(self, nref, params, level_up_max=0):
        """Domain::init_refine with the native vortex-ring refinement condition
        (iblgf_vortex_rings_refine; params as in include/iblgf_gpu.h)."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        fn = ctypes.cast(lib().iblgf_vortex_rings_refine, REFINE_FN)
        _check(lib().iblgf_tree_init_refine(self._h, int(nref), int(level_up_max), fn, p.ctypes.data_as(_P)))
        return self

    def vortex_rings_vorticity(self, params, components=(0, 1, 2)):
        """The listed components of the rings' vorticity on the interior cells of
        every leaf, [len(components), nblocks, n, n, n] (iblgf_vortex_rings_vorticity)."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        nb = self.info()["nblocks"]
        outs = {c: np.zeros((nb, self.n, self.n, self.n)) for c in set(components)}
        mask = sum(1 << c for c in outs)
        ptrs = [outs[c].ctypes.data_as(_DP) if c in outs else None for c in range(3)]
        _check(lib().iblgf_vortex_rings_vorticity(self._h, p.ctypes.data_as(_DP), int(mask), *ptrs))
        return np.stack([outs[c] for c in components])

    def build_masks(self, subtract_non_leaf=False):
        _check(lib().iblgf_tree_build_masks(self._h, 1 if subtract_non_leaf else 0))
        return self

    def adapt(self, changes):
        """Decomposition::adapt_decoposition on one rank (iblgf_tree_adapt):
        changes maps leaf keys to +1 (refine) or -1 (coarsen); returns
        base_mesh_update.  Build the masks again before uploading."""
        keys = np.ascontiguousarray(list(changes.keys()), np.uint64)
        lc = np.ascontiguousarray(list(changes.values()), np.int32)
        upd = ctypes.c_int()
        _check(lib().iblgf_tree_adapt(self._h, len(keys), keys.ctypes.data_as(ctypes.POINTER(_U64)),
                                      lc.ctypes.data_as(_IP), ctypes.byref(upd)))
        return bool(upd.value)

    def info(self):
        nb, bl, dp, be, nl = (ctypes.c_int() for _ in range(5))
        dx = ctypes.c_double()
        _check(lib().iblgf_tree_info(self._h, ctypes.byref(nb), ctypes.byref(bl), ctypes.byref(dp),
                                      ctypes.byref(be), ctypes.byref(dx), ctypes.byref(nl)))
        return dict(nblocks=nb.value, base_level=bl.value, depth=dp.value, block_extent=be.value,
                    dx_base=dx.value, nleaves=nl.value)

    def write_restart(self, path):
        """Tree::write: keys of the octants with data (depth-first) and their leaf flags."""
        _check(lib().iblgf_tree_write_restart(self._h, os.fsencode(path)))

    @classmethod
    def read_restart(cls, path, bd_base, bd_extent, block_extent, dx_base):
        """Domain::initialize_with_keys + restart_list_construct from a Tree::write file."""
        obj = cls.__new__(cls)
        h = ctypes.c_void_p()
        _check(lib().iblgf_tree_read_restart(os.fsencode(path), _ip(bd_base), _ip(bd_extent), int(block_extent),
                                             float(dx_base), ctypes.byref(h)))
        obj._h = h
        obj.block_extent = int(block_extent)
        obj.n = obj.block_extent + 2
        obj.dx_base = float(dx_base)
        return obj

    def write_chombo(self, path, fields, include_correction=True, base_correction_only=True, time=0.0, dt=1.0):
        """H5io::write_h5: fields is a list of (name, array [ncomp, nblocks, n, n, n]) (or [nblocks, n, n, n]
        for one component), written in the Chombo layout."""
        names, ncomp, arrs = [], [], []
        for name, a in fields:
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.ndim == 4:
                a = a[None]
            names.append(name.encode())
            ncomp.append(a.shape[0])
            arrs.append(a)
        nm = (ctypes.c_char_p * len(names))(*names)
        nc = np.ascontiguousarray(ncomp, np.int32)
        dp = (_DP * len(arrs))(*[a.ctypes.data_as(_DP) for a in arrs])
        _check(lib().iblgf_chombo_write(self._h, os.fsencode(path), len(names), nm, nc.ctypes.data_as(_IP), dp,
                                        int(bool(include_correction)), int(bool(base_correction_only)), float(time),
                                        float(dt)))

    def read_chombo(self, path, name, ncomp=1, mesh=0, out=None):
        """H5io::read_h5 (Chombo::read_u): field `name` into [ncomp, nblocks, n, n, n]; blocks outside
        blocks_list_build(true, true) keep the values of `out` (zeros when out is None)."""
        nb = self.info()["nblocks"]
        a = np.zeros((ncomp, nb, self.n, self.n, self.n)) if out is None else np.array(out, dtype=np.float64,
                                                                                        copy=True).reshape(
            ncomp, nb, self.n, self.n, self.n)
        a = np.ascontiguousarray(a)
        _check(lib().iblgf_chombo_read(self._h, os.fsencode(path), name.encode(), int(ncomp), int(mesh),
                                       a.ctypes.data_as(_DP)))
        return a

    def export(self):
        nb = self.info()["nblocks"]
        a = dict(
            key=np.zeros(nb, np.uint64), level=np.zeros(nb, np.int32), coord=np.zeros((nb, 3), np.int32),
            bbase=np.zeros((nb, 3), np.int32), flags=np.zeros(nb, np.uint32), parent=np.zeros(nb, np.int32),
            children=np.zeros((nb, 8), np.int32), neighbor=np.zeros((nb, 27), np.int32),
            influence=np.zeros((nb, 189), np.int32), influence_num=np.zeros(nb, np.int32),
            mask_src=np.zeros(nb, np.uint64), mask_tgt=np.zeros(nb, np.uint64),
        )
        _check(lib().iblgf_tree_export(self._h, *[_ptr(a[k]) for k in (
            "key", "level", "coord", "bbase", "flags", "parent", "children", "neighbor", "influence",
            "influence_num", "mask_src", "mask_tgt")]))
        return a


# ---------------------------------------------------------------------------
# Device solver
# ---------------------------------------------------------------------------
class SolveItem(ctypes.Structure):
    _fields_ = [("kernel", ctypes.c_int), ("alpha_base", ctypes.c_double), ("k", ctypes.c_int),
                ("src_fields", _IP), ("dst_fields", _IP), ("mesh_type", ctypes.c_int)]


class PoissonSolver:
    """PoissonSolver on one GPU; fields are [nblocks, n, n, n] FP64 (x fastest).

    Given a list of trees, the context mirrors them as one forest
    (iblgf_gpu_forest_upload): fields then hold the blocks of every tree in the
    combined slot order, and to_forest / from_forest convert between per-tree
    arrays and forest fields."""

    def __init__(self, tree, device=0):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.iblgf_gpu_create(int(device), ctypes.byref(h)))
        self._h = h
        trees = list(tree) if isinstance(tree, (list, tuple)) else [tree]
        self.trees = trees
        self.tree = trees[0]
        self.n = trees[0].n
        if len(trees) == 1 and not isinstance(tree, (list, tuple)):
            _check(L.iblgf_gpu_tree_upload(self._h, trees[0]._h), self._h)
            self.nblocks = trees[0].info()["nblocks"]
            self.slots = [np.arange(self.nblocks)]
        else:
            arr = (_P * len(trees))(*[t._h for t in trees])
            _check(L.iblgf_gpu_forest_upload(self._h, len(trees), arr), self._h)
            self.slots = []
            for i, t in enumerate(trees):
                sl = np.empty(t.info()["nblocks"], np.int32)
                _check(L.iblgf_gpu_forest_slots(self._h, i, sl.ctypes.data_as(_IP)), self._h)
                self.slots.append(sl)
            self.nblocks = int(sum(len(s) for s in self.slots))
        self.shape = (self.nblocks, self.n, self.n, self.n)

    def readapt(self, tree, fields=(), mesh=MESH_FACE, comp=None):
        """Mirror refresh after tree.adapt + tree.build_masks
        (iblgf_gpu_tree_readapt): fields (one group, field index comp[i],
        default i % 3) carried over by key, new leaves interpolated from their
        parents; returns the number of parents interpolated."""
        fields = [int(f) for f in fields]
        k = len(fields)
        comp = [i % 3 for i in range(k)] if comp is None else list(comp)
        meshes = [int(mesh)] * k if np.isscalar(mesh) else [int(m) for m in mesh]
        npar = ctypes.c_int()
        _check(lib().iblgf_gpu_tree_readapt(self._h, tree._h, k, _ip(fields) if k else None,
                                            _ip(meshes) if k else None, _ip(comp) if k else None,
                                            ctypes.byref(npar)), self._h)
        self.trees = [tree]
        self.tree = tree
        self.nblocks = tree.info()["nblocks"]
        self.slots = [np.arange(self.nblocks)]
        self.shape = (self.nblocks, self.n, self.n, self.n)
        return npar.value

    def coarsen(self, fields, mesh=MESH_FACE):
        """Ifherk::up after clean<F>(true) (iblgf_gpu_coarsen)."""
        _check(lib().iblgf_gpu_coarsen(self._h, _ip(list(fields)), len(fields), int(mesh)), self._h)

    def to_forest(self, arrays):
        """Per-tree arrays [nb_i, n, n, n] -> one forest field array."""
        out = np.zeros(self.shape)
        for sl, a in zip(self.slots, arrays):
            out[sl] = a
        return out

    def from_forest(self, arr):
        """A forest field array -> per-tree arrays in each tree's own slot order."""
        return [arr[sl] for sl in self.slots]

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.iblgf_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def field(self):
        fid = ctypes.c_int()
        _check(lib().iblgf_gpu_field_create(self._h, ctypes.byref(fid)), self._h)
        return fid.value

    def field_f32(self):
        """An FP32 field (iblgf_gpu_field_create_f32); upload/download take float32 arrays."""
        fid = ctypes.c_int()
        _check(lib().iblgf_gpu_field_create_f32(self._h, ctypes.byref(fid)), self._h)
        self._f32 = getattr(self, "_f32", set()) | {fid.value}
        return fid.value

    def device_ptr(self, fid):
        p = ctypes.c_void_p()
        _check(lib().iblgf_gpu_field_device_ptr(self._h, int(fid), ctypes.byref(p)), self._h)
        return p.value

    def upload(self, fid, arr):
        if fid in getattr(self, "_f32", ()):
            a = np.ascontiguousarray(arr, dtype=np.float32)
            if a.size != int(np.prod(self.shape)):
                raise ValueError(f"field has {a.size} values, expected {self.shape}")
            _check(lib().iblgf_gpu_field_upload_f32(self._h, int(fid), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   self._h)
            return
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(self.shape)):
            raise ValueError(f"field has {a.size} values, expected {self.shape}")
        _check(lib().iblgf_gpu_field_upload(self._h, int(fid), a.ctypes.data_as(_DP)), self._h)

    def download(self, fid):
        if fid in getattr(self, "_f32", ()):
            a = np.empty(self.shape, np.float32)
            _check(lib().iblgf_gpu_field_download_f32(self._h, int(fid), a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                   self._h)
            return a
        a = np.empty(self.shape, np.float64)
        _check(lib().iblgf_gpu_field_download(self._h, int(fid), a.ctypes.data_as(_DP)), self._h)
        return a

    def write_chombo(self, path, fields, include_correction=True, base_correction_only=True, time=0.0, dt=1.0):
        """Simulation::write on device fields: fields is a list of (name, [field ids])."""
        names = [n.encode() for n, _ in fields]
        ncomp = np.ascontiguousarray([len(ids) for _, ids in fields], np.int32)
        ids = np.ascontiguousarray([i for _, f in fields for i in f], np.int32)
        nm = (ctypes.c_char_p * len(names))(*names)
        _check(lib().iblgf_gpu_chombo_write(self._h, os.fsencode(path), len(names), nm, ncomp.ctypes.data_as(_IP),
                                            ids.ctypes.data_as(_IP), int(bool(include_correction)),
                                            int(bool(base_correction_only)), float(time), float(dt)), self._h)

    def block_norms(self, fields):
        """Operator::maxnorm and blockRootMeanSquare of the 1-3 component field on every block:
        (maxnorm[nblocks], rms[nblocks])."""
        ids = np.ascontiguousarray(fields, np.int32)
        nb = self.shape[0]
        mx, rms = np.zeros(nb), np.zeros(nb)
        _check(lib().iblgf_gpu_block_norms(self._h, ids.ctypes.data_as(_IP), len(ids), mx.ctypes.data_as(_DP),
                                           rms.ctypes.data_as(_DP)), self._h)
        return mx, rms

    def read_chombo(self, path, name, fields, mesh=0):
        """Simulation::read_h5 into device fields (one id per component)."""
        ids = np.ascontiguousarray(fields, np.int32)
        _check(lib().iblgf_gpu_chombo_read(self._h, os.fsencode(path), name.encode(), len(ids), int(mesh),
                                           ids.ctypes.data_as(_IP)), self._h)

    def set_precision(self, precision):
        """PREC_FP64 (plain); PREC_MIXED (FP32 spectra and MAC products, FP64 DFT legs,
        fields and level sums); PREC_BF16X3, PREC_FP16X3, PREC_FP16, PREC_INT8X4 (as
        PREC_MIXED with the DFT legs on the matrix cores in that arithmetic);
        PREC_FP16X3_FORWARD (forward legs as PREC_FP16X3, inverse legs FP64)."""
        _check(lib().iblgf_gpu_set_precision(self._h, int(precision)), self._h)

    def prepare(self, kernel=KERNEL_LAPLACE, alpha_base=0.0):
        _check(lib().iblgf_gpu_prepare(self._h, int(kernel), float(alpha_base)), self._h)

    def apply_lgf(self, src, dst):
        """PoissonSolver::apply_lgf for one component (asynchronous)."""
        _check(lib().iblgf_gpu_apply_lgf(self._h, int(src), int(dst), KERNEL_LAPLACE, 0.0), self._h)

    def apply_lgf_IF(self, src, dst, alpha_base):
        """PoissonSolver::apply_lgf_IF for one component (asynchronous)."""
        _check(lib().iblgf_gpu_apply_lgf(self._h, int(src), int(dst), KERNEL_IF, float(alpha_base)), self._h)

    def apply_lgf_batch(self, srcs, dsts, kernel=KERNEL_LAPLACE, alpha_base=0.0, mesh=MESH_CELL):
        """The component loop of apply_lgf / apply_lgf_IF (poisson.hpp:130-145,
        419-434) with all components in every launch (asynchronous); mesh is the
        fields' mesh type (MESH_FACE / MESH_EDGE: component c is field index c % 3)."""
        k = len(srcs)
        if len(dsts) != k:
            raise ValueError("srcs and dsts differ in length")
        a = (ctypes.c_int * k)(*[int(x) for x in srcs])
        b = (ctypes.c_int * k)(*[int(x) for x in dsts])
        if mesh == MESH_CELL:
            _check(lib().iblgf_gpu_apply_lgf_batch(self._h, a, b, k, int(kernel), float(alpha_base)), self._h)
        else:
            _check(lib().iblgf_gpu_apply_lgf_mesh(self._h, a, b, k, int(kernel), float(alpha_base), int(mesh)),
                   self._h)

    def set_graphs(self, on):
        """Replay repeated solves from captured HIP graphs (iblgf_gpu_set_graphs)."""
        _check(lib().iblgf_gpu_set_graphs(self._h, 1 if on else 0), self._h)

    def set_concurrency(self, on):
        """Run each FMM pass's coarse chain on a second stream (iblgf_gpu_set_concurrency)."""
        _check(lib().iblgf_gpu_set_concurrency(self._h, 1 if on else 0), self._h)

    def sync(self):
        _check(lib().iblgf_gpu_sync(self._h), self._h)

    def stream(self):
        p = ctypes.c_void_p()
        _check(lib().iblgf_gpu_stream(self._h, ctypes.byref(p)), self._h)
        return p.value

    def counts(self):
        c = (ctypes.c_long * 6)()
        _check(lib().iblgf_gpu_counts(self._h, c), self._h)
        return dict(passes=c[0], n_src=c[1], n_tgt=c[2], n_pair=c[3], n_transfer=c[4], n_leaf=c[5])

    def mac_counts(self):
        """Multiply-accumulate plan statistics of the Laplace passes built so
        far: entries, full-mask entries, pairs in those, all pairs, and the
        partial-mask entries by target count (1..8)."""
        c = (ctypes.c_long * 12)()
        _check(lib().iblgf_gpu_mac_counts(self._h, c), self._h)
        return dict(entries=c[0], full_entries=c[1], full_pairs=c[2], pairs=c[3], masked_by_targets=list(c[4:12]))

    def timing(self, enable):
        _check(lib().iblgf_gpu_timing_enable(self._h, 1 if enable else 0), self._h)

    def timing_reset(self):
        _check(lib().iblgf_gpu_timing_reset(self._h), self._h)

    def timing_read(self):
        """{class: (total_ms, launches, work)} from the HIP events recorded so far."""
        ms = (ctypes.c_double * 5)()
        ln = (ctypes.c_long * 5)()
        wk = (ctypes.c_long * 5)()
        _check(lib().iblgf_gpu_timing_read(self._h, ms, ln, wk), self._h)
        return {k: (ms[i], ln[i], wk[i]) for i, k in enumerate(KERNEL_CLASSES)}

    # -- block operators (domain::Operator) ---------------------------------
    def block_op(self, op, src, dst, level, scale=1.0):
        """dst = op(src) on the interior of the level's blocks (asynchronous); src/dst are field ids
        (an int for cell data, a 3-sequence for face/edge data)."""
        s = [src] if np.isscalar(src) else list(src)
        d = [dst] if np.isscalar(dst) else list(dst)
        _check(lib().iblgf_gpu_block_op(self._h, int(op), _ip(s), _ip(d), int(level), float(scale)), self._h)

    def buffer_exchange(self, field, level):
        _check(lib().iblgf_gpu_buffer_exchange(self._h, int(field), int(level)), self._h)

    def level_divergence(self, src3, dst, level):
        _check(lib().iblgf_gpu_level_divergence(self._h, _ip(src3), int(dst), int(level)), self._h)

    def level_gradient(self, src, dst3, level):
        _check(lib().iblgf_gpu_level_gradient(self._h, int(src), _ip(dst3), int(level)), self._h)

    def clean_boundary(self, field, level, kind, leaf_only=True, width=2):
        _check(lib().iblgf_gpu_clean_boundary(self._h, int(field), int(level), int(kind), 1 if leaf_only else 0,
                                              int(width)), self._h)

    def smooth_boundary(self, field, level, leaf_only):
        """Ifherk::clean_leaf_correction_boundary (iblgf_gpu_smooth_boundary)."""
        _check(lib().iblgf_gpu_smooth_boundary(self._h, int(field), int(level), 1 if leaf_only else 0), self._h)

    def up_and_down(self, fields, mesh=MESH_FACE):
        """Ifherk::up_and_down of the fields (component c is field index c % 3)."""
        _check(lib().iblgf_gpu_up_and_down(self._h, _ip(list(fields)), len(fields), int(mesh)), self._h)

    def nonlinear(self, u3, w3, dst3, level, scale=1.0):
        """Operator::nonlinear on one level: dst = 0.25 (w x u), then times scale."""
        _check(lib().iblgf_gpu_nonlinear(self._h, _ip(u3), _ip(w3), _ip(dst3), int(level), float(scale)), self._h)

    def ifherk_step(self, u3, dt, Re, p=-1, frame=None, refresh=0):
        """One IF-HERK time step (Ifherk::time_step) of the face velocity
        fields u3 in place; p (a field id, or -1) receives the pressure; frame
        is the frame velocity (3 values, or 3 x 3 per stage); refresh 1 opens
        the step with pad_velocity and the adaptation-correction step, 2 with
        pad_velocity only."""
        fr = None
        if frame is not None:
            fr = np.ascontiguousarray(np.broadcast_to(np.asarray(frame, np.float64), (3, 3)))
        _check(lib().iblgf_gpu_ifherk_step(self._h, _ip(u3), int(p), float(dt), float(Re),
                                           None if fr is None else fr.ctypes.data_as(_DP), int(refresh)), self._h)

    def apply_lgf_typed(self, src, dst, fmm_type=FMM_AMR2AMR, kernel=KERNEL_LAPLACE, alpha_base=0.0):
        """apply_lgf / apply_lgf_IF with the reference's fmm_type
        (poisson.hpp:130-135, 427-433); types without a device path raise
        IblgfError with code IBLGF_ERR_UNSUPPORTED (4)."""
        _check(lib().iblgf_gpu_apply_lgf_typed(self._h, int(src), int(dst), int(kernel), float(alpha_base),
                                               int(fmm_type)), self._h)

    def solve_sequence(self, items):
        """items: (kernel, alpha_base, srcs, dsts) tuples, issued in order as
        apply_lgf_batch calls; with graphs on, one captured graph for all."""
        keep = []
        arr = (SolveItem * len(items))()
        for i, item in enumerate(items):
            kernel, alpha, srcs, dsts = item[:4]
            mesh = item[4] if len(item) > 4 else MESH_CELL
            a = (ctypes.c_int * len(srcs))(*srcs)
            b = (ctypes.c_int * len(dsts))(*dsts)
            keep += [a, b]
            arr[i] = SolveItem(int(kernel), float(alpha), len(srcs), ctypes.cast(a, _IP), ctypes.cast(b, _IP),
                               int(mesh))
        _check(lib().iblgf_gpu_solve_sequence(self._h, len(items), ctypes.cast(arr, _P)), self._h)

    def autotune(self, srcs, dsts, kernel=KERNEL_LAPLACE, alpha_base=0.0, reps=3):
        """Time the batched solve under every result-preserving launch shape and
        keep the fastest (iblgf_gpu_autotune); returns (report, choice)."""
        a = (ctypes.c_int * len(srcs))(*srcs)
        b = (ctypes.c_int * len(dsts))(*dsts)
        buf = ctypes.create_string_buffer(4096)
        _check(lib().iblgf_gpu_autotune(self._h, ctypes.cast(a, _IP), ctypes.cast(b, _IP), len(srcs), int(kernel),
                                        float(alpha_base), int(reps), buf, len(buf)), self._h)
        return buf.value.decode(), self.launch_shapes()

    def set_launch_shapes(self, mac_policy, dft_grid):
        _check(lib().iblgf_gpu_set_launch_shapes(self._h, int(mac_policy), int(dft_grid)), self._h)

    def launch_shapes(self):
        m, d, s = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().iblgf_gpu_launch_shapes(self._h, ctypes.byref(m), ctypes.byref(d), ctypes.byref(s)), self._h)
        return {"mac_policy": m.value, "dft_grid": d.value, "streams": s.value}

    def apply_lgf_noalloc(self, src, dst, kernel=KERNEL_LAPLACE, alpha_base=0.0):
        """apply_lgf / apply_lgf_IF of one component with its work memory in the
        library's fixed per-GPU scratch (no device allocation during the call);
        prepare(kernel, alpha_base) first, in the current precision."""
        _check(lib().iblgf_gpu_apply_lgf_noalloc(self._h, int(src), int(dst), int(kernel), float(alpha_base)),
               self._h)

    def apply_lgf_masked(self, src, dst, mask_type, kernel=KERNEL_LAPLACE, alpha_base=0.0):
        """apply_lgf / apply_lgf_IF with fmm_type IB2xIB or xIB2IB (finest level, one pass)."""
        _check(lib().iblgf_gpu_apply_lgf_masked(self._h, int(src), int(dst), int(kernel), float(alpha_base),
                                                int(mask_type)), self._h)


class IbSystem:
    """Immersed-boundary force system on the device (LinSysSolver on one GPU).

    Forces are (N, 3) FP64 host arrays; the points and delta kernel come from
    the tree (Tree.set_ib), which must be the one the solver uploaded."""

    def __init__(self, solver):
        self.solver = solver
        _check(lib().iblgf_gpu_ib_create(solver._h, solver.tree._h), solver._h)
        npts, nown, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        _check(lib().iblgf_gpu_ib_info(solver._h, ctypes.byref(npts), ctypes.byref(nown), ctypes.byref(nnz)),
               solver._h)
        self.npoints, self.nowned, self.nnz = npts.value, nown.value, nnz.value

    def _vec(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        if a.size != 3 * self.npoints:
            raise ValueError(f"force has {a.size} values, expected {3 * self.npoints}")
        return a

    def smear(self, force, fields3):
        _check(lib().iblgf_gpu_ib_smear(self.solver._h, self._vec(force).ctypes.data_as(_DP), _ip(fields3)),
               self.solver._h)

    def project(self, fields3):
        out = np.zeros((self.npoints, 3), np.float64)
        _check(lib().iblgf_gpu_ib_project(self.solver._h, _ip(fields3), out.ctypes.data_as(_DP)), self.solver._h)
        return out

    def apply(self, fin, alpha):
        """A fin with A = E (I - G L^-1 D) H_alpha E^T (ET_H_S_E)."""
        out = np.zeros((self.npoints, 3), np.float64)
        _check(lib().iblgf_gpu_ib_apply(self.solver._h, self._vec(fin).ctypes.data_as(_DP),
                                        out.ctypes.data_as(_DP), float(alpha)), self.solver._h)
        return out

    def ifherk_step(self, u3, dt, Re, p=-1, frame=None, refresh=0, bc=None, cg_tol=1e-3, cg_maxit=100):
        """One IF-HERK step with these IB points (iblgf_gpu_ifherk_step_ib);
        returns the three main stages' CG iteration counts."""
        def arr9(x):
            return None if x is None else np.ascontiguousarray(
                np.broadcast_to(np.asarray(x, np.float64), (3, 3))).ctypes.data_as(_DP)

        fr, b = arr9(frame), arr9(bc)
        its = (ctypes.c_int * 3)()
        _check(lib().iblgf_gpu_ifherk_step_ib(self.solver._h, _ip(u3), int(p), float(dt), float(Re), fr,
                                              int(refresh), b, float(cg_tol), int(cg_maxit), its), self.solver._h)
        return list(its)

    def set_boundary_velocity(self, velocity):
        """Per-point boundary velocities [3 stages, N, 3] for the following IF-HERK steps
        (LinSysSolver::subtract_boundary_vel); None returns to the steps' constant bc."""
        if velocity is None:
            _check(lib().iblgf_gpu_ib_set_boundary_velocity(self.solver._h, None), self.solver._h)
            return
        a = np.ascontiguousarray(velocity, np.float64).reshape(-1)
        if a.size != 9 * self.npoints:
            raise ValueError("boundary velocities must hold 3 x N x 3 values")
        _check(lib().iblgf_gpu_ib_set_boundary_velocity(self.solver._h, a.ctypes.data_as(_DP)), self.solver._h)

    def stage_forces(self, value=None):
        """The IF-HERK stage forces [3, N, 3] (set them when value is given)."""
        if value is not None:
            a = np.ascontiguousarray(value, np.float64).reshape(-1)
            if a.size != 9 * self.npoints:
                raise ValueError("stage forces must hold 3 x N x 3 values")
            _check(lib().iblgf_gpu_ib_stage_forces(self.solver._h, a.ctypes.data_as(_DP), 1), self.solver._h)
        out = np.zeros((3, self.npoints, 3))
        _check(lib().iblgf_gpu_ib_stage_forces(self.solver._h, out.ctypes.data_as(_DP), 0), self.solver._h)
        return out

    def cg(self, uc, alpha, tol=1e-8, maxit=200, f0=None):
        """CG_solve; returns (f, iterations, residual, history)."""
        f = np.zeros(3 * self.npoints) if f0 is None else self._vec(f0).copy()
        it, res = ctypes.c_int(), ctypes.c_double()
        hist = np.zeros(max(int(maxit), 1))
        _check(lib().iblgf_gpu_ib_cg(self.solver._h, self._vec(uc).ctypes.data_as(_DP), f.ctypes.data_as(_DP),
                                     float(alpha), float(tol), int(maxit), ctypes.byref(it), ctypes.byref(res),
                                     hist.ctypes.data_as(_DP)), self.solver._h)
        return f.reshape(-1, 3), it.value, res.value, hist[: it.value].copy()

    def fcg(self, uc, alpha, tol=1e-8, maxit=200, f0=None, use_mlp=True):
        """Flexible CG in conjugate-residual form with the learned preconditioner (identity when
        use_mlp is False); returns (f, it, res, history)."""
        f = np.zeros(3 * self.npoints) if f0 is None else self._vec(f0).copy()
        it, res = ctypes.c_int(), ctypes.c_double()
        hist = np.zeros(max(int(maxit), 1))
        _check(lib().iblgf_gpu_ib_fcg(self.solver._h, self._vec(uc).ctypes.data_as(_DP), f.ctypes.data_as(_DP),
                                      float(alpha), float(tol), int(maxit), 1 if use_mlp else 0, ctypes.byref(it),
                                      ctypes.byref(res), hist.ctypes.data_as(_DP)), self.solver._h)
        return f.reshape(-1, 3), it.value, res.value, hist[: it.value].copy()

    STOP = ("converged", "iteration limit", "breakdown")

    def stats(self):
        """(seconds, operator applications, preconditioner fell back, stop reason) of the last solve."""
        sec, ap, fb, st = ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib().iblgf_gpu_ib_stats(self.solver._h, ctypes.byref(sec), ctypes.byref(ap), ctypes.byref(fb),
                                        ctypes.byref(st)), self.solver._h)
        return sec.value, ap.value, bool(fb.value), self.STOP[st.value]

    def recycle(self, capacity):
        """Keep up to `capacity` (p, A p) pairs across flexible solves with the same alpha (0: off)."""
        _check(lib().iblgf_gpu_ib_recycle(self.solver._h, int(capacity)), self.solver._h)

    def switched_at(self):
        """Iteration after which the last learned solve continued with the identity (-1: never)."""
        it = ctypes.c_int()
        _check(lib().iblgf_gpu_ib_switch_iteration(self.solver._h, ctypes.byref(it)), self.solver._h)
        return it.value


class LearnedPreconditioner:
    """The MLP preconditioner of an IbSystem's flexible CG (trained on the device)."""

    def __init__(self, system, K=16, hidden=128, seed=7):
        self.system = system
        h = system.solver._h
        self._h = h
        _check(lib().iblgf_gpu_ib_precond_create(h, system.solver.tree._h, int(K), int(hidden), int(seed)), h)
        k, hid, dim, npar, cs = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long(), ctypes.c_double()
        _check(lib().iblgf_gpu_ib_precond_info(h, ctypes.byref(k), ctypes.byref(hid), ctypes.byref(dim),
                                               ctypes.byref(npar), ctypes.byref(cs)), h)
        self.K, self.hidden, self.dim, self.nparam = k.value, hid.value, dim.value, npar.value

    def collect(self, stride=1):
        """Record every stride-th iterate of the following solves (0 stops)."""
        _check(lib().iblgf_gpu_ib_precond_collect(self._h, int(stride)), self._h)

    def train(self, steps=2000, batch_rows=4096, lr=1e-3):
        loss, npairs = ctypes.c_double(), ctypes.c_int()
        _check(lib().iblgf_gpu_ib_precond_train(self._h, int(steps), int(batch_rows), float(lr), ctypes.byref(loss),
                                                ctypes.byref(npairs)), self._h)
        return loss.value, npairs.value

    def output_scale(self):
        cs = ctypes.c_double()
        _check(lib().iblgf_gpu_ib_precond_info(self._h, None, None, None, None, ctypes.byref(cs)), self._h)
        return cs.value

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        _check(lib().iblgf_gpu_ib_precond_params(self._h, _ptr(p), self.nparam), self._h)
        return p

    def set_params(self, p):
        q = np.ascontiguousarray(p, np.float32)
        _check(lib().iblgf_gpu_ib_precond_set_params(self._h, _ptr(q), q.size), self._h)

    def features(self, r):
        X = np.zeros((self.system.npoints, self.dim), np.float32)
        _check(lib().iblgf_gpu_ib_precond_features(self._h, self.system._vec(r).ctypes.data_as(_DP), _ptr(X)),
               self._h)
        return X

    def loss_and_grad(self, X, T):
        x = np.ascontiguousarray(X, np.float32)
        t = np.ascontiguousarray(T, np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_double()
        _check(lib().iblgf_gpu_ib_precond_grad(self._h, _ptr(x), _ptr(t), x.shape[0], _ptr(g), ctypes.byref(loss)),
               self._h)
        return loss.value, g


def lgf_get(x, y, z):
    return lib().iblgf_lgf_get(int(x), int(y), int(z))


def lgf_table(nmax=100):
    g = np.zeros((nmax + 1) * (nmax + 2) * (nmax + 3) // 6, np.float64)
    _check(lib().iblgf_lgf_table(int(nmax), g.ctypes.data_as(_DP)))
    return g


def if_kernel(x, y, z, alpha):
    return lib().iblgf_if_kernel(int(x), int(y), int(z), float(alpha))


def lagrange_matrix(n, pts_cap, relative_pos):
    A = np.zeros((n, 2 * n), np.float64)
    _check(lib().iblgf_lagrange_matrix(int(n), int(pts_cap), int(relative_pos), A.ctypes.data_as(_DP)))
    return A


# ---------------------------------------------------------------------------
# Screened 2-D LGFs of the spanwise Fourier modes (device arrays from torch)
# ---------------------------------------------------------------------------
def _device_f64(t, numel, what):
    if not (t.is_cuda and t.dtype.is_floating_point and t.element_size() == 8 and t.is_contiguous()):
        raise ValueError(f"{what}: a contiguous float64 tensor on the GPU")
    if t.numel() < numel:
        raise ValueError(f"{what}: needs {numel} values, has {t.numel()}")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def helm2d_table(c, nx, ny, table):
    """Fill table (torch float64 [nmodes, ny, nx] on the GPU) with the screened
    2-D LGF values of screenings c (one per Fourier mode) on torch's current
    stream (Helmholtz::get for Dim = 2, include/iblgf/lgf/helmholtz.hpp:143-153)."""
    c = np.ascontiguousarray(np.atleast_1d(c), np.float64)
    tp = _device_f64(table, c.size * nx * ny, "helm2d_table")
    _check(lib().iblgf_gpu_helm2d_table(c.ctypes.data_as(_DP), int(c.size), int(nx), int(ny), tp, _stream()))


def helm2d_convolve(table, src, dst, scale=1.0):
    """dst = scale * (free-space convolution of src with table) per mode, all
    torch float64 [nmodes, ny, nx] on the GPU, on torch's current stream."""
    nmodes, ny, nx = src.shape
    n = nmodes * ny * nx
    _check(lib().iblgf_gpu_helm2d_convolve(_device_f64(table, n, "table"), int(nmodes), int(nx), int(ny),
                                           _device_f64(src, n, "src"), _device_f64(dst, n, "dst"), float(scale),
                                           _stream()))
