"""The C ABI library loads, exports every symbol include/iblgf_gpu.h declares,
and reports argument and state errors with status codes (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

import iblgf_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "iblgf_gpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iblgf_[a-z0-9_]+)\s*\(", text)) - {"iblgf_refine_fn"})


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name
    # the Python binding types every declared symbol
    assert set(names) <= set(P.SYMBOLS)


def test_argument_errors_return_status():
    L = P.lib()
    out = ctypes.c_void_p()
    ext = (ctypes.c_int * 3)(32, 32, 32)
    base = (ctypes.c_int * 3)(0, 0, 0)
    rc = L.iblgf_tree_create(base, ext, 0, 1.0, 0, None, ctypes.byref(out))
    assert rc == 1 and b"invalid" in L.iblgf_last_error()
    ext_bad = (ctypes.c_int * 3)(30, 32, 32)
    regs = (ctypes.c_int * 6)(0, 0, 0, 32, 32, 32)
    rc = L.iblgf_tree_create(base, ext_bad, 8, 1.0, 1, regs, ctypes.byref(out))
    assert rc == 1 and b"divisible" in L.iblgf_last_error()
    regs_bad = (ctypes.c_int * 6)(0, 0, 0, 20, 32, 32)
    rc = L.iblgf_tree_create(base, ext, 8, 1.0, 1, regs_bad, ctypes.byref(out))
    assert rc != 0 and b"divisible" in L.iblgf_last_error()
    assert L.iblgf_tree_build_masks(None, 0) == 1
    assert L.iblgf_gpu_apply_lgf(None, 0, 1, 0, 0.0) == 1
    assert L.iblgf_gpu_create(0, None) == 1
    assert L.iblgf_lagrange_matrix(18, 5, 3, None) == 1


def test_refinement_callback_errors_propagate():
    t = P.Tree((0, 0, 0), (64, 64, 64), 8, 1.0, [(16, 16, 16, 32, 32, 32)])

    def bad(base, ext, rl, diff):
        raise ValueError("callback failure")

    with pytest.raises(ValueError):
        t.init_refine(1, bad)


def test_device_creation_without_gpu_fails_loudly():
    """On a host without a GPU the device path refuses to start (no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    t = P.Tree((0, 0, 0), (32, 32, 32), 8, 1.0, [(0, 0, 0, 32, 32, 32)]).init_refine(0).build_masks()
    with pytest.raises(P.IblgfError) as e:
        P.PoissonSolver(t)
    assert e.value.code == 3  # IBLGF_ERR_DEVICE


def test_tree_export_shapes():
    t = P.Tree((-16, -16, -16), (32, 32, 32), 8, 1 / 32, [(-16, -16, -16, 32, 32, 32)]).init_refine(0).build_masks()
    a = t.export()
    nb = t.info()["nblocks"]
    assert a["neighbor"].shape == (nb, 27) and a["influence"].shape == (nb, 189)
    assert np.all(a["key"] & np.uint64(1) == 1)


def build_c_caller(tmp_path):
    """tests/capi_check.c compiled as C99 against include/iblgf_gpu.h and
    linked to the built library."""
    import subprocess

    exe = str(tmp_path / "capi_check")
    libdir = os.path.dirname(P.LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "capi_check.c"), "-o", exe, "-L", libdir, "-liblgf_gpu",
           f"-Wl,-rpath,{libdir}"]
    subprocess.run(cmd, check=True)
    return exe


def test_c_caller_status_codes(tmp_path):
    """A C caller sees ERR_UNSUPPORTED for the fmm types without a device
    path (Laplacian_BC, Fourier modes, STREAM or IB2AMR with the IF kernel) and for a tree whose
    masks were built with subtract_non_leaf = 1, ERR_ARG otherwise."""
    import subprocess

    r = subprocess.run([build_c_caller(tmp_path), "cpu"], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def test_subtract_non_leaf_refused_from_python():
    t = P.Tree((-16, -16, -16), (32, 32, 32), 8, 1 / 32, [(-16, -16, -16, 32, 32, 32)]).init_refine(0)
    t.build_masks(subtract_non_leaf=True)
    assert t.export()["mask_src"].any()  # the masks themselves are built and exported
    L = P.lib()
    assert L.iblgf_gpu_tree_upload(None, t._h) == 4  # IBLGF_ERR_UNSUPPORTED
    assert b"subtract_non_leaf" in L.iblgf_last_error()


def test_noalloc_entry_refuses_null_context():
    L = P.lib()
    assert L.iblgf_gpu_apply_lgf_noalloc(None, 0, 1, 0, 0.0) == 1
    assert b"null context" in L.iblgf_last_error()
