"""The native vortex-ring workload (iblgf_vortex_rings_refine and
iblgf_vortex_rings_vorticity) against its numpy restatement: the same tree
(keys and flags bit for bit) and the same vorticity to rounding."""
import numpy as np

from iblgf_amd import Tree, setups


def _small():
    return setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                              radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)


def test_native_refinement_builds_the_numpy_tree():
    v = _small()
    region = 4 * 8
    half = region // 2
    t = Tree((-region,) * 3, (2 * region,) * 3, 8, v.dx_base, [(-half, -half, -half, region, region, region)])
    t.init_refine(2, v._refine)  # the numpy condition through a Python callback
    t.build_masks()
    a = t.export()
    b = v.arrays
    assert a["key"].shape == b["key"].shape and np.array_equal(a["key"], b["key"])
    assert np.array_equal(a["flags"], b["flags"])


def test_native_vorticity_equals_numpy():
    v = _small()
    got = v.sources()
    ref = v.sources_numpy()
    scale = np.max(np.abs(ref))
    assert scale > 0
    assert np.max(np.abs(got - ref)) < 1e-13 * scale
    leaves = (v.arrays["flags"] & 1) != 0
    # zero on non-leaf blocks and on every block's buffer ring
    assert not np.any(got[:, ~leaves])
    assert not np.any(got[:, :, 0]) and not np.any(got[:, :, -1])
    assert np.array_equal(v.source(1), got[1])
