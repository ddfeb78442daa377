"""Per-block norms on the device (iblgf_gpu_block_norms: Operator::maxnorm
and blockRootMeanSquare, operators.hpp:1149-1167, 1351-1367) and the
reference setup's adaptation criterion evaluated from them
(setups.adapt_level_changes, tests/ns_amr_lgf/ns_amr_lgf.hpp:402-495): the
largest norms equal the host's from the downloaded fields bit for bit, the
root mean squares to 1e-14, so the level changes are the host's; the tree
adapts with them and the device mirror follows (needs an MI355X)."""
import numpy as np
import pytest

from iblgf_amd import MESH_FACE, PoissonSolver, setups

pytestmark = pytest.mark.gpu


def _host_norms(fields):
    x = [f[:, 1:-1, 1:-1, 1:-1] for f in fields]
    tmp = x[0] * x[0]
    for v in x[1:]:
        tmp = tmp + v * v
    nb = tmp.shape[0]
    return np.sqrt(tmp.reshape(nb, -1).max(axis=1)), np.sqrt(tmp.reshape(nb, -1).mean(axis=1))


def test_block_norms_and_adaptation_criterion():
    print("case: block norms of 1- and 3-component fields on the small rings tree; ns_amr_lgf criterion", flush=True)
    v = setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                           radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)
    t, a, info = v.tree, v.arrays, v.info
    s = PoissonSolver(t)
    w = [v.source(c) for c in range(3)]  # the rings' vorticity: edge_aux of the criterion
    d = 0.01 * w[0]  # a scalar field of the rings' support: cell_aux
    fw = [s.field() for _ in range(3)]
    fd = s.field()
    for f, x in zip(fw, w):
        s.upload(f, x)
    s.upload(fd, d)
    mx3, rms3 = s.block_norms(fw)
    hx3, hr3 = _host_norms(w)
    assert np.array_equal(mx3, hx3)
    assert np.allclose(rms3, hr3, rtol=1e-14, atol=0)
    mx1, rms1 = s.block_norms([fd])
    hx1, hr1 = _host_norms([d])
    assert np.array_equal(mx1, hx1)
    assert np.allclose(rms1, hr1, rtol=1e-14, atol=0)

    source_max = (float(rms1.max()), float(rms3.max()))  # Ifherk::update_source_max
    norms = {"cell_aux": mx1, "correction_tmp": mx1, "edge_aux": mx3}
    host = {"cell_aux": hx1, "correction_tmp": hx1, "edge_aux": hx3}
    ch = setups.adapt_level_changes(a, info["base_level"], norms, source_max, nref=2, refinement_factor=0.5)
    assert ch == setups.adapt_level_changes(a, info["base_level"], host, (float(hr1.max()), float(hr3.max())),
                                            nref=2, refinement_factor=0.5)
    assert any(c > 0 for c in ch.values()) and any(c < 0 for c in ch.values())
    print(f"{sum(c > 0 for c in ch.values())} refine, {sum(c < 0 for c in ch.values())} coarsen", flush=True)
    nb0 = info["nblocks"]
    t.adapt(ch)
    t.build_masks()
    s.readapt(t, fields=fw, mesh=MESH_FACE)
    assert t.info()["nblocks"] != nb0 and s.shape[0] == t.info()["nblocks"]
    mx_after, _ = s.block_norms(fw)
    assert np.all(np.isfinite(mx_after)) and mx_after.shape[0] == s.shape[0]
