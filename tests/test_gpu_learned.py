"""Learned preconditioner of the IB force CG on the device (needs an MI355X).

* The training kernels' loss and gradients are checked against the FP64
  restatement oracle/mlp_ref.py with the same BF16-rounded weights
  (FP32 accumulation on the device: relative L2 <= 1e-4).
* Features: the rotated self residual keeps the residual's norm and the
  self offset is zero.
* Training on recorded solve pairs lowers the loss; the learned flexible
  solve is consistent (finite, true residual = recurrence residual) and its
  iteration gain over the identity is measured.
"""
import numpy as np
import pytest

from iblgf_amd import IbSystem, LearnedPreconditioner, PoissonSolver, setups
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    print("case: IB sphere N=600, block extent 8, 2 refinements", flush=True)
    s = setups.IbSphere(npoints=600, block_extent=8, nref=2, half_width=1.0)
    solver = PoissonSolver(s.tree)
    sysm = IbSystem(solver)
    return s, solver, sysm


def test_training_gradients_match_fp64_reference(small):
    s, solver, sysm = small
    pc = LearnedPreconditioner(sysm, K=16, hidden=128, seed=3)
    rng = np.random.default_rng(21)
    rows = 257
    p = (rng.standard_normal(pc.nparam) * 0.2).astype(np.float32)
    pc.set_params(p)
    X = rng.standard_normal((rows, pc.dim)).astype(np.float32)
    T = rng.standard_normal((rows, 3)).astype(np.float32)
    loss, g = pc.loss_and_grad(X, T)
    lref, gref = mlp_ref.loss_and_grad(p.astype(np.float64), X, T, pc.dim, pc.hidden, bf16_weights=True)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"loss {loss:.6e} vs {lref:.6e}; grad relL2 {rel:.3e}")
    assert abs(loss - lref) < 1e-5 * abs(lref) and rel < 1e-4
    assert np.array_equal(pc.params(), p)


def test_feature_invariants(small):
    s, solver, sysm = small
    pc = LearnedPreconditioner(sysm, K=16, hidden=128, seed=3)
    rng = np.random.default_rng(22)
    r = rng.standard_normal((s.npoints, 3))
    X = pc.features(r)
    sc = np.sqrt(np.mean(r ** 2))
    assert np.allclose(np.linalg.norm(X[:, :3], axis=1) * sc, np.linalg.norm(r, axis=1), rtol=1e-5)
    assert np.all(X[:, 3 * pc.K:3 * pc.K + 3] == 0.0)


def test_learned_preconditioner_trains_and_solves(small):
    """Training on the device lowers the loss on pairs recorded from solves,
    and the flexible conjugate-residual solver with the network returns a
    finite force whose true residual agrees with its recurrence.  The learned
    solve is monitored against the last identity solve's convergence rate and
    hands over to the identity when the network is slower, so it converges
    within a bounded margin of the identity's iteration count.  The gain over
    the identity is measured and printed; on this operator (half its spectrum
    below 1e-8, rough surface modes) the local network does not reduce the
    iteration count, which the design notes record."""
    s, solver, sysm = small
    alpha, tol = 0.5, 1e-5
    pc = LearnedPreconditioner(sysm, K=16, hidden=128, seed=5)
    pc.collect(stride=2)
    rng = np.random.default_rng(23)
    for t in range(4):
        b = rng.standard_normal((s.npoints, 3)) if t % 2 else np.tile(rng.standard_normal(3), (s.npoints, 1))
        sysm.fcg(b, alpha, tol=1e-6, maxit=800, use_mlp=False)
    pc.collect(0)
    loss0, npairs = pc.train(steps=10, batch_rows=600, lr=1e-3)
    loss1, _ = pc.train(steps=2000, batch_rows=600, lr=1e-3)
    print(f"pairs {npairs} loss {loss0:.4e} -> {loss1:.4e}", flush=True)
    assert npairs > 0 and loss1 < loss0
    f_id, it_id, res_id, _ = sysm.fcg(s.uc, alpha, tol=tol, maxit=1500, use_mlp=False)
    print(f"identity iterations {it_id} residual {res_id:.3e} stop {sysm.stats()[3]}", flush=True)
    assert sysm.stats()[3] == "converged"
    f_l, it_l, res_l, _ = sysm.fcg(s.uc, alpha, tol=tol, maxit=1500)
    assert np.all(np.isfinite(f_l))
    true_res = np.linalg.norm(s.uc - sysm.apply(f_l, alpha)) / np.linalg.norm(f_l)
    print(f"learned iterations {it_l} residual {res_l:.3e} true {true_res:.3e}; identity {it_id}; "
          f"gain {it_id / max(it_l, 1):.2f}; stop {sysm.stats()[3]} fell_back {sysm.stats()[2]} "
          f"identity from iteration {sysm.switched_at()}")
    assert abs(true_res - res_l) <= 1e-3 * max(true_res, 1e-12) + 1e-12
    assert sysm.stats()[3] == "converged" and it_l <= 1.5 * it_id + 20


def test_krylov_recycling_cuts_iterations(small):
    """The flexible solver keeps the (p, A p) pairs of earlier solves with the
    same alpha; a later solve of another right-hand side, deflated by them,
    needs fewer iterations than the same solve without them and its true
    residual agrees with the recurrence.  A solve with another alpha drops
    the pairs instead of using them."""
    s, solver, sysm = small
    alpha, tol = 0.5, 1e-5
    rng = np.random.default_rng(41)
    sysm.recycle(0)
    f0, it0, res0, _ = sysm.fcg(s.uc, alpha, tol=tol, maxit=1500, use_mlp=False)
    assert sysm.stats()[3] == "converged"
    sysm.recycle(150)
    for t in range(2):
        b = rng.standard_normal((s.npoints, 3)) if t else np.tile(rng.standard_normal(3), (s.npoints, 1))
        sysm.fcg(b, alpha, tol=1e-6, maxit=800, use_mlp=False)
    f1, it1, res1, _ = sysm.fcg(s.uc, alpha, tol=tol, maxit=1500, use_mlp=False)
    true_res = np.linalg.norm(s.uc - sysm.apply(f1, alpha)) / np.linalg.norm(f1)
    print(f"without recycling {it0} iterations; with 150 recycled pairs {it1} "
          f"(residual {res1:.3e}, true {true_res:.3e}); gain {it0 / max(it1, 1):.2f}")
    assert sysm.stats()[3] == "converged" and it1 < it0
    assert abs(true_res - res1) <= 1e-3 * max(true_res, 1e-12) + 1e-12
    f2, it2, res2, _ = sysm.fcg(s.uc, 0.3, tol=tol, maxit=1500, use_mlp=False)
    assert sysm.stats()[3] == "converged" and np.all(np.isfinite(f2))
    sysm.recycle(0)


def test_non_finite_network_falls_back_to_the_identity(small):
    """A network whose parameters are NaN (set from the host, so every kernel
    still runs to completion) makes every preconditioned direction
    non-finite; the flexible solve must take the residual itself instead,
    report the fallback, and converge exactly as the identity solve does."""
    s, solver, sysm = small
    alpha, tol = 0.5, 1e-5
    pc = LearnedPreconditioner(sysm, K=16, hidden=128, seed=9)
    p = pc.params()
    p[:] = np.nan
    pc.set_params(p)
    print("case: NaN network parameters, flexible solve on the 600-point sphere", flush=True)
    sysm.recycle(0)
    f_id, it_id, res_id, _ = sysm.fcg(s.uc, alpha, tol=tol, maxit=1500, use_mlp=False)
    f_l, it_l, res_l, _ = sysm.fcg(s.uc, alpha, tol=tol, maxit=1500, use_mlp=True)
    st = sysm.stats()
    print(f"identity {it_id} iterations; NaN network {it_l} iterations, fell back {st[2]}, stop {st[3]}", flush=True)
    assert st[2] and st[3] == "converged"
    assert np.all(np.isfinite(f_l))
    assert it_l == it_id
    assert np.allclose(f_l, f_id, rtol=1e-9, atol=1e-12 * np.max(np.abs(f_id)))
