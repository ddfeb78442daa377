"""The adaptive vortex-ring workload at full size (8^3 base blocks of 16^3
cells, 3 refinement levels, 16 seeded Gaussian-core rings: 4290 blocks, n =
18) against the FP64 oracle.

Tolerances, as everywhere: the plain FP64 path agrees with the oracle to
relative L2 <= 1e-10 (same arithmetic, different summation order); the
batched mixed path (FP32 spectra storage and FP32 products, FP64 DFT legs and
level sums) to relative L2 <= 1e-6 per field component, the precision target
of the solve.  The oracle runs multi-threaded (OpenMP over targets; its
results do not depend on the thread count).
"""
import time

import numpy as np
import pytest

from iblgf_amd import FLAG_LEAF, KERNEL_IF, PREC_FP16X3_FORWARD, PREC_MIXED, PoissonSolver, setups

pytestmark = pytest.mark.gpu

_STATE = {}


def _rel(a, b):
    d = a - b
    return np.linalg.norm(d) / np.linalg.norm(b), np.max(np.abs(d)) / np.max(np.abs(b))


def _workload(oracle_mod):
    if not _STATE:
        t0 = time.time()
        v = setups.VortexRings()
        srcs = v.sources()
        print(f"workload: {v.info} ({time.time() - t0:.1f} s)", flush=True)
        _STATE.update(v=v, srcs=srcs, ot=oracle_mod.Tree(v.arrays, v.info, v.tree.n), ref={})
    return _STATE


def _ref(st, c):
    if c not in st["ref"]:
        t0 = time.time()
        st["ref"][c] = st["ot"].apply_lgf(st["srcs"][c])
        print(f"oracle component {c}: {time.time() - t0:.1f} s", flush=True)
    return st["ref"][c]


def test_full_size_fp64_matches_oracle(oracle_mod):
    st = _workload(oracle_mod)
    v, srcs = st["v"], st["srcs"]
    print(f"case: full vortex-ring tree, FP64, component 0, {v.info['nblocks']} blocks", flush=True)
    assert v.tree.n == 18 and v.info["nblocks"] > 4000
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, srcs[0])
    solver.apply_lgf(fs, fd)
    got = solver.download(fd)
    ref = _ref(st, 0)
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    l2, mx = _rel(got[leaves], ref[leaves])
    print(f"FP64 vs oracle: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
    assert l2 < 1e-10 and mx < 1e-9


def test_full_size_mixed_batched_matches_oracle(oracle_mod):
    """The benchmarked path (three components batched, mixed precision) on
    every component, and plain FP64 batched equal to FP64 single solves."""
    st = _workload(oracle_mod)
    v, srcs = st["v"], st["srcs"]
    print("case: full vortex-ring tree, mixed batched, 3 components", flush=True)
    solver = PoissonSolver(v.tree)
    fin = [solver.field() for _ in range(3)]
    fout = [solver.field() for _ in range(3)]
    for f, x in zip(fin, srcs):
        solver.upload(f, x)
    solver.set_precision(PREC_MIXED)
    solver.apply_lgf_batch(fin, fout)
    got = [solver.download(f) for f in fout]
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    for c in range(3):
        l2, mx = _rel(got[c][leaves], _ref(st, c)[leaves])
        print(f"mixed batched component {c}: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
        assert l2 < 1e-6


def test_full_size_fp16x3_forward_batched_matches_oracle(oracle_mod):
    """The matrix-core forward legs (FP16 three-term) with FP64 inverse legs,
    three components batched: within relative L2 1e-6 of the oracle on every
    component of the full tree."""
    st = _workload(oracle_mod)
    v, srcs = st["v"], st["srcs"]
    print("case: full vortex-ring tree, fp16x3 forward legs, batched, 3 components", flush=True)
    solver = PoissonSolver(v.tree)
    fin = [solver.field() for _ in range(3)]
    fout = [solver.field() for _ in range(3)]
    for f, x in zip(fin, srcs):
        solver.upload(f, x)
    solver.set_precision(PREC_FP16X3_FORWARD)
    solver.apply_lgf_batch(fin, fout)
    got = [solver.download(f) for f in fout]
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    for c in range(3):
        l2, mx = _rel(got[c][leaves], _ref(st, c)[leaves])
        print(f"fp16x3 forward batched component {c}: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
        assert l2 < 1e-6


def test_full_size_if_matches_oracle(oracle_mod):
    """The integrating factor at a stage alpha of the IF-HERK scheme, FP64,
    against the oracle on every block (apply_lgf_IF writes every block)."""
    st = _workload(oracle_mod)
    v, srcs = st["v"], st["srcs"]
    alpha = 0.7
    print(f"case: full vortex-ring tree, IF alpha={alpha}", flush=True)
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, srcs[2])
    solver.apply_lgf_batch([fs], [fd], kernel=KERNEL_IF, alpha_base=alpha)
    got = solver.download(fd)
    t0 = time.time()
    ref = st["ot"].apply_if(srcs[2], alpha)
    print(f"oracle IF: {time.time() - t0:.1f} s", flush=True)
    l2, mx = _rel(got, ref)
    print(f"IF vs oracle: relL2={l2:.3e} relmax={mx:.3e}", flush=True)
    assert l2 < 1e-10 and mx < 1e-9
