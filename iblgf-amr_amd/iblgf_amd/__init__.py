"""Python binding of libiblgf_gpu.so, the MI355X LGF fast-multipole path.

The classes mirror the reference's interfaces for this path:

* :class:`Tree` -- ``Domain`` / ``Tree`` construction, ``init_refine``,
  ``FmmMaskBuilder`` (include/iblgf/domain/domain.hpp:106-211, 878-943;
  include/iblgf/fmm/fmm.hpp:392-680);
* :class:`PoissonSolver` -- ``PoissonSolver::apply_lgf`` / ``apply_lgf_IF``
  (include/iblgf/solver/poisson/poisson.hpp:130-135, 427-494, 587-894) on one
  GPU with the tree resident.

Everything calls the C ABI declared in ``include/iblgf_gpu.h``.  There is no
CPU fallback: constructing a :class:`PoissonSolver` without the built library
or without a GPU raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libiblgf_gpu.so")

KERNEL_LAPLACE = 0
KERNEL_IF = 1

FLAG_LEAF = 1
FLAG_CORRECTION = 2
FLAG_LEAF_BOUNDARY = 4
FLAG_PHYSICAL = 8

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
    "iblgf_tree_build_masks": (_I, [_P, _I]),
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
    "iblgf_gpu_prepare": (_I, [_P, _I, _D]),
    "iblgf_gpu_sync": (_I, [_P]),
    "iblgf_gpu_stream": (_I, [_P, ctypes.POINTER(_P)]),
    "iblgf_gpu_counts": (_I, [_P, ctypes.POINTER(ctypes.c_long)]),
    "iblgf_gpu_timing_enable": (_I, [_P, _I]),
    "iblgf_gpu_timing_read": (_I, [_P, _DP, ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long)]),
    "iblgf_gpu_timing_reset": (_I, [_P]),
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

    def init_refine(self, nref, cond=None, level_up_max=0):
        """Domain::init_refine.  cond(block_base(3,), extent, refinement_level, diff_level) -> bool."""
        err = []

        def _cb(user, base, ext, rl, diff):
            try:
                return 1 if cond((base[0], base[1], base[2]), ext, rl, diff) else 0
            except Exception as e:  # noqa: BLE001 - surfaced after the call
                err.append(e)
                return 0

        cb = REFINE_FN(_cb)
        _check(lib().iblgf_tree_init_refine(self._h, int(nref), int(level_up_max), cb, None))
        if err:
            raise err[0]
        return self

    def build_masks(self, subtract_non_leaf=False):
        _check(lib().iblgf_tree_build_masks(self._h, 1 if subtract_non_leaf else 0))
        return self

    def info(self):
        nb, bl, dp, be, nl = (ctypes.c_int() for _ in range(5))
        dx = ctypes.c_double()
        _check(lib().iblgf_tree_info(self._h, ctypes.byref(nb), ctypes.byref(bl), ctypes.byref(dp),
                                      ctypes.byref(be), ctypes.byref(dx), ctypes.byref(nl)))
        return dict(nblocks=nb.value, base_level=bl.value, depth=dp.value, block_extent=be.value,
                    dx_base=dx.value, nleaves=nl.value)

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
class PoissonSolver:
    """PoissonSolver on one GPU; fields are [nblocks, n, n, n] FP64 (x fastest)."""

    def __init__(self, tree, device=0):
        L = lib()
        h = ctypes.c_void_p()
        _check(L.iblgf_gpu_create(int(device), ctypes.byref(h)))
        self._h = h
        self.tree = tree
        _check(L.iblgf_gpu_tree_upload(self._h, tree._h), self._h)
        info = tree.info()
        self.nblocks = info["nblocks"]
        self.n = tree.n
        self.shape = (self.nblocks, self.n, self.n, self.n)

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

    def device_ptr(self, fid):
        p = ctypes.c_void_p()
        _check(lib().iblgf_gpu_field_device_ptr(self._h, int(fid), ctypes.byref(p)), self._h)
        return p.value

    def upload(self, fid, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != int(np.prod(self.shape)):
            raise ValueError(f"field has {a.size} values, expected {self.shape}")
        _check(lib().iblgf_gpu_field_upload(self._h, int(fid), a.ctypes.data_as(_DP)), self._h)

    def download(self, fid):
        a = np.empty(self.shape, np.float64)
        _check(lib().iblgf_gpu_field_download(self._h, int(fid), a.ctypes.data_as(_DP)), self._h)
        return a

    def prepare(self, kernel=KERNEL_LAPLACE, alpha_base=0.0):
        _check(lib().iblgf_gpu_prepare(self._h, int(kernel), float(alpha_base)), self._h)

    def apply_lgf(self, src, dst):
        """PoissonSolver::apply_lgf for one component (asynchronous)."""
        _check(lib().iblgf_gpu_apply_lgf(self._h, int(src), int(dst), KERNEL_LAPLACE, 0.0), self._h)

    def apply_lgf_IF(self, src, dst, alpha_base):
        """PoissonSolver::apply_lgf_IF for one component (asynchronous)."""
        _check(lib().iblgf_gpu_apply_lgf(self._h, int(src), int(dst), KERNEL_IF, float(alpha_base)), self._h)

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
