"""The reference's Poisson acceptance test on the device: the manufactured
vortex ring of tests/poisson/vortexrings.hpp with the L-infinity bounds of
tests/poisson/configs/config_1_lv_0:51 and config_1_lv_1:51, through the C
ABI, plus parity with the oracle on the same trees (relative L2 <= 1e-10 in
FP64; the mixed path must meet the same analytic bounds).
"""
import numpy as np
import pytest

from iblgf_amd import FLAG_LEAF, PREC_MIXED, PoissonSolver, setups

pytestmark = pytest.mark.gpu

BOUNDS = {0: 0.0116450, 1: 1e-2}


@pytest.mark.parametrize("nlevels,rf", [(0, 0.125), (1, 0.125), (1, 0.6)])
def test_vortex_ring_known_answer(oracle_mod, nlevels, rf):
    print(f"case: analytic vortex ring, {nlevels} refinement levels, refinement factor {rf}", flush=True)
    v = setups.AnalyticVortexRing(nlevels=nlevels, refinement_factor=rf)
    src = v.source()
    solver = PoissonSolver(v.tree)
    fs, fd = solver.field(), solver.field()
    solver.upload(fs, src)
    solver.apply_lgf(fs, fd)
    got = solver.download(fd)
    err = v.linf_error(got)
    ref = oracle_mod.Tree(v.arrays, v.info, v.tree.n).apply_lgf(src)
    leaves = (v.arrays["flags"] & FLAG_LEAF) != 0
    d = got[leaves] - ref[leaves]
    l2 = np.linalg.norm(d) / np.linalg.norm(ref[leaves])
    print(f"{v.info['nblocks']} blocks: Linf={err:.6g} (oracle {v.linf_error(ref):.6g}), relL2 vs oracle={l2:.3e}",
          flush=True)
    assert l2 < 1e-10
    assert err < BOUNDS[nlevels]
    solver.set_precision(PREC_MIXED)
    solver.apply_lgf(fs, fd)
    errm = v.linf_error(solver.download(fd))
    print(f"mixed: Linf={errm:.6g}", flush=True)
    assert errm < BOUNDS[nlevels]
