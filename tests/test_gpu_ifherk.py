"""The IF-HERK time step on the device (iblgf_gpu_ifherk_step) against the
oracle's restatement (ifherk.hpp:1078-1200, no immersed-boundary points).

The field operations of the step are the reference's operation sequences
with FP64 roundings in the same order, so the nonlinear term and the ramped
boundary clean equal the oracle bit for bit; the level transfers of
up_and_down sum in another order (relative 1e-13); the step itself goes
through six LGF and up to twelve integrating-factor component solves, which
agree with the oracle to relative L2 <= 1e-10, and the step is held to that.
"""
import numpy as np
import pytest

from iblgf_amd import FLAG_LEAF, MESH_EDGE, MESH_FACE, PoissonSolver, setups

pytestmark = pytest.mark.gpu

BLOBS = [((0.02, -0.03, 0.01), 0.09, (0, 0, 1), 1.0), ((-0.04, 0.02, 0.03), 0.08, (1, 0.5, 0), 0.7),
         ((0.03, 0.04, -0.02), 0.085, (0.2, 1, 0.4), -0.8)]


def _rings():
    return setups.VortexRings(seed=7, block_extent=8, base_blocks=4, nref=2, nrings=3, centre_range=4.0,
                              radius=(4.0, 8.0), core=(1.5, 2.5), refinement_factor=0.5)


def _rel(a, b):
    return np.sqrt(sum(np.sum((x - y) ** 2) for x, y in zip(a, b))) / np.sqrt(sum(np.sum(y ** 2) for y in b))


def _fields(solver):
    return [solver.field() for _ in range(3)]


def test_field_operations_match_oracle(oracle_mod):
    print("case: nonlinear, smooth boundary, up_and_down on the small adaptive rings tree (n = 10)", flush=True)
    v = _rings()
    n, info = v.tree.n, v.info
    ot = oracle_mod.Tree(v.arrays, info, n)
    rng = np.random.default_rng(5)
    shape = (info["nblocks"], n, n, n)
    u = [rng.standard_normal(shape) for _ in range(3)]
    w = [rng.standard_normal(shape) for _ in range(3)]
    s = PoissonSolver(v.tree)
    fu, fw, fd = _fields(s), _fields(s), _fields(s)
    for f, x in zip(fu + fw, u + w):
        s.upload(f, x)
    for l in range(info["base_level"], info["depth"]):
        for f in fd:
            s.upload(f, np.zeros(shape))
        s.nonlinear(fu, fw, fd, l, scale=-0.37)
        got = [s.download(f) for f in fd]
        ref = ot.nonlinear(u, w, l, scale=-0.37)
        assert all(np.array_equal(g, r) for g, r in zip(got, ref)), l
    for leaf_only in (False, True):
        for l in (info["base_level"], info["base_level"] + 1):
            s.upload(fd[0], u[0])
            s.smooth_boundary(fd[0], l, leaf_only)
            ref = u[0].copy()
            ot.smooth_boundary(ref, l, leaf_only)
            assert np.array_equal(s.download(fd[0]), ref), (leaf_only, l)
    for mesh in (MESH_FACE, MESH_EDGE):
        for f, x in zip(fd, u):
            s.upload(f, x)
        s.up_and_down(fd, mesh)
        got = [s.download(f) for f in fd]
        ref = ot.up_and_down(u, mesh)
        err = _rel(got, ref)
        print(f"up_and_down mesh {mesh}: relL2 {err:.2e}", flush=True)
        assert err < 1e-13


@pytest.mark.parametrize("case", ["adaptive", "adaptive_frame_inviscid", "uniform", "adaptive_refresh",
                                  "adaptive_refresh_pad_only"])
def test_step_matches_oracle(oracle_mod, case):
    print(f"case: IF-HERK step, {case}", flush=True)
    if case == "uniform":
        g = setups.UniformGaussians(nblocks_side=4, block_extent=8)
        tree = g.tree
    else:
        tree = _rings().tree
    flow = setups.GaussianFlow(tree, BLOBS)
    u0 = flow.velocity()
    info, n = flow.info, tree.n
    dt, Re, frame, refresh = 2e-4, 500.0, None, 0
    if case == "adaptive_frame_inviscid":
        Re, frame = np.inf, ((0.3, -0.1, 0.2), (0.31, -0.1, 0.2), (0.33, -0.1, 0.2))
    if case.startswith("adaptive_refresh"):
        refresh = 2 if case.endswith("pad_only") else 1
    ot = oracle_mod.Tree(flow.arrays, info, n)
    ref_u, ref_p = ot.ifherk_step(u0, dt, Re, frame, refresh=refresh)
    s = PoissonSolver(tree)
    fu = [s.field() for _ in range(3)]
    fp = s.field()
    for f, x in zip(fu, u0):
        s.upload(f, x)
    s.ifherk_step(fu, dt, Re, p=fp, frame=frame, refresh=refresh)
    got_u = [s.download(f) for f in fu]
    got_p = s.download(fp)
    leaves = (flow.arrays["flags"] & FLAG_LEAF) != 0
    eu = _rel([x[leaves] for x in got_u], [x[leaves] for x in ref_u])
    ep = _rel([got_p[leaves]], [ref_p[leaves]])
    change = _rel([x[leaves] for x in ref_u], [x[leaves] for x in u0])
    print(f"velocity relL2 {eu:.2e}, pressure relL2 {ep:.2e} (step changed u by {change:.2e})", flush=True)
    assert 1e-4 < change < 0.1
    assert eu < 1e-10 and ep < 1e-10
    # every block, not only the leaves
    assert _rel(got_u, ref_u) < 1e-10


def test_step_graph_replays_equal_eager():
    print("case: IF-HERK steps (one opening with the velocity refresh) captured as graphs, replayed three times",
          flush=True)
    tree = _rings().tree
    flow = setups.GaussianFlow(tree, BLOBS)
    u0 = flow.velocity()
    s = PoissonSolver(tree)
    fu = [s.field() for _ in range(3)]
    fp = s.field()
    for f, x in zip(fu, u0):
        s.upload(f, x)
    s.ifherk_step(fu, 2e-4, 500.0, p=fp, refresh=1)
    s.ifherk_step(fu, 2e-4, 500.0, p=fp)
    ref = [s.download(f) for f in fu] + [s.download(fp)]
    s.set_graphs(True)
    for rep in range(3):
        for f, x in zip(fu, u0):
            s.upload(f, x)
        s.ifherk_step(fu, 2e-4, 500.0, p=fp, refresh=1)
        s.ifherk_step(fu, 2e-4, 500.0, p=fp)
        s.sync()
        got = [s.download(f) for f in fu] + [s.download(fp)]
        assert all(np.array_equal(g, r) for g, r in zip(got, ref)), rep
        s.apply_lgf(fu[0], fp)  # eager work between replays
    s.set_graphs(False)


def test_step_refuses_bad_arguments():
    from iblgf_amd import IBLGF_ERR_ARG, IblgfError

    tree = _rings().tree
    s = PoissonSolver(tree)
    fu = [s.field() for _ in range(3)]
    for bad in (dict(dt=0.0, Re=1.0), dict(dt=1e-3, Re=-1.0)):
        with pytest.raises(IblgfError) as e:
            s.ifherk_step(fu, bad["dt"], bad["Re"])
        assert e.value.code == IBLGF_ERR_ARG
    with pytest.raises(IblgfError):
        s.ifherk_step([fu[0], fu[0], fu[1]], 1e-3, 1.0)
    with pytest.raises(IblgfError):
        s.ifherk_step(fu, 1e-3, 1.0, p=fu[1])
    with pytest.raises(IblgfError) as e:
        s.ifherk_step(fu, 1e-3, 1.0, refresh=3)
    assert e.value.code == IBLGF_ERR_ARG


def test_ib_step_matches_oracle_and_enforces_the_boundary_velocity(oracle_mod):
    """iblgf_gpu_ifherk_step_ib against the oracle's step with the same IB
    points: the force solves converge (tolerance 1e-10) to the same forces and
    velocities (relative 1e-7: the CG iterates differ in rounding, not in the
    solution), the velocity at the owned points is the boundary velocity, and
    the stage forces persist into the next step (its initial guesses)."""
    from iblgf_amd import IbSystem

    print("case: IF-HERK step with 100 IB points on the sphere mesh (n = 10)", flush=True)
    sph = setups.IbSphere(npoints=100, block_extent=8, nref=1, half_width=0.5)
    flow = setups.GaussianFlow(sph.tree, [((0.1, 0.0, 0.0), 0.3, (0, 0, 1), 0.5)])
    u0 = flow.velocity()
    dt = 0.1 * sph.dx_base / max(np.abs(x).max() for x in u0)
    bc = np.array([0.2, -0.1, 0.05])
    ot = oracle_mod.Tree(sph.arrays, sph.info, sph.tree.n)
    oib = oracle_mod.Ib(sph.points, sph.nref, sph.delta, 4.5, sph.safety_dis)
    fp = np.zeros((3, sph.npoints, 3))
    ref_u, ref_p = oib.ifherk_step(ot, u0, dt, 200.0, fp, bc=bc, cg_tol=1e-10, cg_maxit=400)
    s = PoissonSolver(sph.tree)
    ib = IbSystem(s)
    fu = [s.field() for _ in range(3)]
    fpr = s.field()
    for f, x in zip(fu, u0):
        s.upload(f, x)
    its1 = ib.ifherk_step(fu, dt, 200.0, p=fpr, bc=bc, cg_tol=1e-10, cg_maxit=400)
    got_u = [s.download(f) for f in fu]
    forces = ib.stage_forces()
    leaves = (sph.arrays["flags"] & FLAG_LEAF) != 0
    eu = _rel([x[leaves] for x in got_u], [x[leaves] for x in ref_u])
    ef = np.linalg.norm(forces - fp) / np.linalg.norm(fp)
    own = oib.owned(ot).astype(bool)
    slip = np.abs(ib.project(fu)[own] - bc).max()
    print(f"velocity relL2 {eu:.2e}, stage forces relL2 {ef:.2e}, slip {slip:.2e}, CG iterations {its1}", flush=True)
    assert eu < 1e-7 and ef < 1e-6 and slip < 1e-8
    # the stage forces persist: a second step from them equals the oracle's
    # second step from its kept forces
    ref_u2, _ = oib.ifherk_step(ot, ref_u, dt, 200.0, fp, bc=bc, cg_tol=1e-10, cg_maxit=400)
    its2 = ib.ifherk_step(fu, dt, 200.0, p=fpr, bc=bc, cg_tol=1e-10, cg_maxit=400)
    eu2 = _rel([s.download(f)[leaves] for f in fu], [x[leaves] for x in ref_u2])
    ef2 = np.linalg.norm(ib.stage_forces() - fp) / np.linalg.norm(fp)
    print(f"second step: CG iterations {its2}, velocity relL2 {eu2:.2e}, stage forces relL2 {ef2:.2e}", flush=True)
    assert eu2 < 1e-7 and ef2 < 1e-6
    # set forces round-trip
    ib.stage_forces(np.zeros((3, sph.npoints, 3)))
    assert not ib.stage_forces().any()
    from iblgf_amd import IBLGF_ERR_STATE, IblgfError
    s2 = PoissonSolver(sph.tree)
    g = [s2.field() for _ in range(3)]
    from iblgf_amd import lib
    import ctypes
    a = (ctypes.c_int * 3)(*g)
    assert lib().iblgf_gpu_ifherk_step_ib(s2._h, a, -1, dt, 200.0, None, 0, None, 1e-6, 10, None) == IBLGF_ERR_STATE


def test_ib_step_with_per_point_boundary_velocities(oracle_mod):
    """iblgf_gpu_ib_set_boundary_velocity (subtract_boundary_vel point by
    point, linsys.hpp:148-158) for a translating, spinning body: the device
    step equals the oracle's (velocity relative 1e-7, stage forces 1e-6), the
    velocity at the owned points is the last stage's prescribed velocity, and
    the same velocity at every point gives the constant-bc step bit for bit."""
    from iblgf_amd import IbSystem

    print("case: IF-HERK step with 100 IB points, per-point boundary velocities of a spinning body", flush=True)
    sph = setups.IbSphere(npoints=100, block_extent=8, nref=1, half_width=0.5)
    flow = setups.GaussianFlow(sph.tree, [((0.1, 0.0, 0.0), 0.3, (0, 0, 1), 0.5)])
    u0 = flow.velocity()
    dt = 0.1 * sph.dx_base / max(np.abs(x).max() for x in u0)
    vp = []
    for t in (dt / 3, dt, dt):
        om = 0.8 + 0.3 * t
        v = np.zeros_like(sph.points)
        v[:, 0], v[:, 1], v[:, 2] = 0.1 - om * sph.points[:, 1], -0.05 + om * sph.points[:, 0], 0.02
        vp.append(v)
    vp = np.stack(vp)
    ot = oracle_mod.Tree(sph.arrays, sph.info, sph.tree.n)
    oib = oracle_mod.Ib(sph.points, sph.nref, sph.delta, 4.5, sph.safety_dis)
    fp = np.zeros((3, sph.npoints, 3))
    ref_u, _ = oib.ifherk_step(ot, u0, dt, 200.0, fp, cg_tol=1e-10, cg_maxit=400, bc_points=vp)
    s = PoissonSolver(sph.tree)
    ib = IbSystem(s)
    fu = [s.field() for _ in range(3)]
    fpr = s.field()
    for f, x in zip(fu, u0):
        s.upload(f, x)
    ib.set_boundary_velocity(vp)
    ib.ifherk_step(fu, dt, 200.0, p=fpr, cg_tol=1e-10, cg_maxit=400)
    leaves = (sph.arrays["flags"] & FLAG_LEAF) != 0
    eu = _rel([s.download(f)[leaves] for f in fu], [x[leaves] for x in ref_u])
    ef = np.linalg.norm(ib.stage_forces() - fp) / np.linalg.norm(fp)
    own = oib.owned(ot).astype(bool)
    slip = np.abs(ib.project(fu)[own] - vp[2][own]).max()
    print(f"velocity relL2 {eu:.2e}, stage forces relL2 {ef:.2e}, slip {slip:.2e}", flush=True)
    assert eu < 1e-7 and ef < 1e-6 and slip < 1e-8
    # the same velocity at every point is the constant-bc step
    bc = np.array([[0.2, -0.1, 0.05], [0.1, 0.0, -0.2], [0.3, 0.1, 0.0]])
    outs = []
    for per_point in (False, True):
        for f, x in zip(fu, u0):
            s.upload(f, x)
        ib.stage_forces(np.zeros((3, sph.npoints, 3)))
        ib.set_boundary_velocity(np.repeat(bc[:, None, :], sph.npoints, axis=1) if per_point else None)
        ib.ifherk_step(fu, dt, 200.0, p=fpr, bc=None if per_point else bc, cg_tol=1e-10, cg_maxit=400)
        outs.append([s.download(f) for f in fu] + [s.download(fpr), ib.stage_forces()])
    assert all(np.array_equal(a, b) for a, b in zip(*outs))


@pytest.mark.parametrize("config", [0, 1])
def test_device_ifherk_matches_the_reference_vortex_ring_run(oracle_mod, config):
    """The reference's known answer (tests/ns_amr_lgf with ref.hdf5, fixture
    tests/golden/ns_amr_lgf_ref.npz) on the device: the ring initialised by
    the batched edge-mesh LGF solve and -curl^T per level, four
    iblgf_gpu_ifherk_step calls, the base level cleaned; u_0 on the finest
    non-correction leaves within the reference's EXP_LInf of its u_0, and the
    whole velocity within 1e-10 of the oracle's run."""
    import os

    from iblgf_amd import MESH_EDGE, OP_CURL_TRANSPOSE

    print(f"case: ns_amr_lgf configFile_{config} on the device", flush=True)
    ns = setups.NsAmrLgfRing(config)
    ref = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ns_amr_lgf_ref.npz")))
    info = ns.info
    L = info["base_level"]
    s = PoissonSolver(ns.tree)
    w = [s.field() for _ in range(3)]
    psi = [s.field() for _ in range(3)]
    u = [s.field() for _ in range(3)]
    for f, x in zip(w, ns.vorticity()):
        s.upload(f, x)
    s.apply_lgf_batch(w, psi, mesh=MESH_EDGE)
    for l in range(L, info["depth"]):
        s.block_op(OP_CURL_TRANSPOSE, psi, u, l, scale=-1.0)
        for c in range(3):
            s.buffer_exchange(u[c], l)
    for _ in range(ns.STEPS):
        s.ifherk_step(u, ns.DT, ns.RE)
    for c in range(3):
        s.smooth_boundary(u[c], L, True)
    got = [s.download(f) for f in u]
    err = ns.reference_error(got[0], ref)
    # the oracle's run
    import test_ifherk_oracle as T

    ot = oracle_mod.Tree(ns.arrays, info, ns.tree.n)
    v = T._ns_init_oracle(ot, ns)
    for _ in range(ns.STEPS):
        v, _p = ot.ifherk_step(v, ns.DT, ns.RE)
    for c in range(3):
        ot.smooth_boundary(v[c], L, True)
    leaves = (ns.arrays["flags"] & FLAG_LEAF) != 0
    e_or = _rel([x[leaves] for x in got], [x[leaves] for x in v])
    print(f"config {config}: L_inf vs ref.hdf5 {err:.3e} (EXP_LInf {ns.EXP_LINF}), relL2 vs oracle {e_or:.2e}",
          flush=True)
    assert err < ns.EXP_LINF and e_or < 1e-10
