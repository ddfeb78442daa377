"""The allocation-free entry point iblgf_gpu_apply_lgf_noalloc, through the C
ABI: its results equal iblgf_gpu_apply_lgf's bit for bit (same kernels, same
order, only the work memory differs) on two calls in a row, the second on a
smaller tree, for the Laplace and integrating-factor kernels and in mixed
precision; it refuses unprepared solves (IBLGF_ERR_STATE) and trees whose work
exceeds the fixed scratch (IBLGF_ERR_UNSUPPORTED)."""
import numpy as np
import pytest

from iblgf_amd import (IBLGF_ERR_STATE, IBLGF_ERR_UNSUPPORTED, KERNEL_IF, KERNEL_LAPLACE, PREC_MIXED, IblgfError,
                       PoissonSolver, setups)

pytestmark = pytest.mark.gpu


def _rings(extent):
    return setups.VortexRings(seed=7, block_extent=extent, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                              radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)


def _both(solver, src, kernel=KERNEL_LAPLACE, alpha=0.0):
    fs, fa, fb = solver.field(), solver.field(), solver.field()
    solver.upload(fs, src)
    # both destinations start as NaN: every value either call writes must be
    # written, and the blocks a solve leaves alone stay NaN in both
    solver.upload(fa, np.full_like(src, np.nan))
    solver.upload(fb, np.full_like(src, np.nan))
    solver.prepare(kernel, alpha)
    solver.apply_lgf_noalloc(fs, fb, kernel=kernel, alpha_base=alpha)
    if kernel == KERNEL_LAPLACE:
        solver.apply_lgf(fs, fa)
    else:
        solver.apply_lgf_IF(fs, fa, alpha)
    return solver.download(fa), solver.download(fb)


def _equal_where_written(a, b, kernel, flags):
    from iblgf_amd import FLAG_LEAF

    assert np.array_equal(a, b, equal_nan=True)
    # the values the reference's caller reads are all written
    sel = np.ones(len(flags), bool) if kernel == KERNEL_IF else (flags & FLAG_LEAF) != 0
    if kernel == KERNEL_LAPLACE:
        assert np.all(np.isfinite(b[sel]))


def test_noalloc_equals_apply_lgf_on_two_trees_in_a_row():
    big = _rings(6)  # 514 blocks, n = 8
    small = setups.UniformGaussians(seed=1000)  # 74 blocks, n = 10
    for name, wl in (("rings n=8", big), ("uniform n=10", small)):
        print(f"case: noalloc vs apply_lgf, {name}, {wl.info['nblocks']} blocks", flush=True)
        solver = PoissonSolver(wl.tree)
        src = wl.source(0) if hasattr(wl, "sources") else wl.source()
        fs, fd = solver.field(), solver.field()
        solver.upload(fs, src)
        with pytest.raises(IblgfError) as e:
            solver.apply_lgf_noalloc(fs, fd)  # nothing prepared yet
        assert e.value.code == IBLGF_ERR_STATE
        flags = wl.arrays["flags"]
        a, b = _both(solver, src)
        _equal_where_written(a, b, KERNEL_LAPLACE, flags)
        a, b = _both(solver, src, KERNEL_IF, 0.4)
        _equal_where_written(a, b, KERNEL_IF, flags)


def test_noalloc_mixed_precision_and_size_limit():
    wl = _rings(8)  # 514 blocks, n = 10: 72 MiB of FP64 work, 46 MiB in mixed precision
    print(f"case: noalloc size limit and mixed precision, {wl.info['nblocks']} blocks n=10", flush=True)
    solver = PoissonSolver(wl.tree)
    src = wl.source(1)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.prepare()
    with pytest.raises(IblgfError) as e:
        solver.apply_lgf_noalloc(fs, fd)
    assert e.value.code == IBLGF_ERR_UNSUPPORTED
    solver.set_precision(PREC_MIXED)
    a, b = _both(solver, src)
    _equal_where_written(a, b, KERNEL_LAPLACE, wl.arrays["flags"])
