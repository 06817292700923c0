"""Plans: a batched shift + solve + refinement step captured in a HIP graph
and replayed for new shifts, with the refinement decided on the device.
Every run is checked against the eager batched call on the same shifts (the
per-column FP64 residual <= 1e-12 for both, the solutions within 1e-12 n of
each other), across the learning run, the capture and the replays, with the
program's own eager work and synchronizations between the replays; a batch
that needs more rounds than the graph holds continues round by round and is
captured again; a tolerance that cannot be met is reported."""
import numpy as np
import pytest

from inputs import randn, stiff

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def col_resid(J, X, B, sig, tau, offs):
    r = 0.0
    for s in range(len(sig)):
        c = slice(offs[s], offs[s + 1])
        W = sig[s] * np.eye(J.shape[0]) + tau[s] * J
        r = max(r, np.abs(B[:, c] - W @ X[:, c]).max() / np.abs(B[:, c]).max())
    return r


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_plan_replays_match_eager(lhl_hip, dtype):
    import torch
    n, S, R = 700, 8, 4
    J = stiff(61, n)
    rng = np.random.default_rng(3)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    Bh = np.asfortranarray(rng.standard_normal((n, S * R)))
    B = torch.from_numpy(Bh).cuda()
    X = torch.zeros_like(B)
    ws = lhl_hip.lhl(J, dtype=dtype)
    plan = lhl_hip.BatchPlan(ws, offs, B, X, precision="refine64", tol=1e-12, max_iter=30)
    for run, (lo, hi) in enumerate(((-4, -1), (-4, -1.2), (-3.5, -1), (-4, -1.1), (-4.2, -1.3))):
        gam = np.logspace(lo, hi, S)
        say(f"case: plan run {run} {np.dtype(dtype).name} gamma [{gam[0]:.2g}, {gam[-1]:.2g}]")
        info, iters = plan.run(np.ones(S), -gam)
        Xp = X.cpu().numpy()
        Xe, info_e, it_e = lhl_hip.shift_solve_batched(ws, np.ones(S), -gam, Bh, offs, precision="refine64",
                                                       tol=1e-12, max_iter=30)
        assert np.array_equal(info, info_e)
        rp = col_resid(J, Xp, Bh, np.ones(S), -gam, offs)
        assert rp <= 1e-12, (run, rp)
        e = np.linalg.norm(Xp - Xe) / np.linalg.norm(Xe)
        assert e <= 1e-12 * n, (run, e)
        say(f"  rounds {iters.min()}-{iters.max()} (eager {it_e.min()}-{it_e.max()}), resid {rp:.2e}, "
            f"vs eager {e:.2e}, stats {plan.stats()}")
        # the program's own eager work between replays (a reduction of another J
        # on another workspace, a single solve on this one)
        ws2 = lhl_hip.lhl(randn(62 + run, 300))
        lhl_hip.lhl_shift_(ws, 1.0, -0.01)
        x = lhl_hip.lhl_ldiv_(Bh[:, 0].copy(), ws)
        torch.cuda.synchronize()
        del ws2, x
    g, replays, captures, extra = plan.stats()
    assert g >= 1 and replays >= 3 and captures >= 1
    plan.close()


def test_plan_complex_shifts_and_new_B(lhl_hip):
    import torch
    n, S, R = 500, 6, 3
    rng = np.random.default_rng(4)
    A = np.asfortranarray(rng.standard_normal((n, n)) / np.sqrt(n) - 2 * np.eye(n))
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    Bh = np.asfortranarray(rng.standard_normal((n, S * R)) + 1j * rng.standard_normal((n, S * R)))
    B = torch.from_numpy(Bh).cuda()
    X = torch.zeros_like(B)
    ws = lhl_hip.lhl(A, dtype=np.float32)
    plan = lhl_hip.BatchPlan(ws, offs, B, X, shifts_complex=True, tol=1e-12, max_iter=30)
    for run in range(4):
        om = np.logspace(-2 + 0.1 * run, 2, S)
        if run == 3:                                  # a new right-hand side, same buffer
            Bh = np.asfortranarray(rng.standard_normal((n, S * R)) + 1j * rng.standard_normal((n, S * R)))
            B.copy_(torch.from_numpy(Bh))
        say(f"case: complex plan run {run}")
        info, iters = plan.run(1j * om, -np.ones(S))
        assert np.all(info == 0)
        Xp = X.cpu().numpy()
        rp = col_resid(A, Xp, Bh, 1j * om, -np.ones(S), offs)
        assert rp <= 1e-12, (run, rp)
        Xe, _, _ = lhl_hip.shift_solve_batched(ws, 1j * om, -np.ones(S), Bh, offs, precision="refine64", tol=1e-12,
                                               max_iter=30)
        assert np.linalg.norm(Xp - Xe) / np.linalg.norm(Xe) <= 1e-12 * n
    assert plan.stats()[1] >= 2


def test_plan_more_rounds_than_graph(lhl_hip):
    """Well-conditioned shifts first (few rounds captured), then shifts next to
    an eigenvalue of J: the replay continues round by round and the next run
    captures the larger count."""
    import torch
    n, S, R = 600, 4, 2
    rng = np.random.default_rng(65)
    J = np.asfortranarray(rng.standard_normal((n, n)) / np.sqrt(n) - 2 * np.eye(n))
    lam = np.linalg.eigvals(J)
    lr = lam[np.abs(lam.imag) < 1e-12].real
    lam0 = lr[np.argmax(lr)]                        # a real eigenvalue of J
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    Bh = np.asfortranarray(rng.standard_normal((n, S * R)))
    B = torch.from_numpy(Bh).cuda()
    X = torch.zeros_like(B)
    ws = lhl_hip.lhl(J, dtype=np.float32)
    plan = lhl_hip.BatchPlan(ws, offs, B, X, tol=1e-12, max_iter=30)
    easy = np.full(S, 10.0)                          # W = J + 10 I
    hard = -lam0 + np.logspace(-3, -2, S)            # W = J - (lam0 - delta) I, near singular
    rounds = []
    for run, sg in enumerate((easy, easy, easy, hard, hard)):
        say(f"case: rounds plan run {run}")
        info, iters = plan.run(sg, np.ones(S))
        rp = col_resid(J, X.cpu().numpy(), Bh, sg, np.ones(S), offs)
        assert rp <= 1e-12, (run, rp)
        rounds.append(int(iters.max()))
    gr, replays, captures, extra = plan.stats()
    say(f"  rounds per run {rounds}, stats {plan.stats()}")
    assert rounds[3] > rounds[2]
    assert extra >= 1 and captures >= 2 and gr >= rounds[3]


def test_plan_not_converged_and_working(lhl_hip):
    import torch
    n, S, R = 300, 3, 2
    J = stiff(64, n)
    offs = np.arange(0, S * R + 1, R, dtype=np.int64)
    Bh = np.asfortranarray(np.random.default_rng(6).standard_normal((n, S * R)))
    B = torch.from_numpy(Bh).cuda()
    X = torch.zeros_like(B)
    ws = lhl_hip.lhl(J, dtype=np.float32)
    gam = np.logspace(-3, -1, S)
    plan = lhl_hip.BatchPlan(ws, offs, B, X, tol=1e-12, max_iter=0)
    for run in range(3):
        with pytest.raises(lhl_hip.LHLNotConverged):
            plan.run(np.ones(S), -gam)
    # no refinement: the working-precision solve, against the eager call's
    ws64 = lhl_hip.lhl(J)
    plan2 = lhl_hip.BatchPlan(ws64, offs, B, X, precision="working")
    for run in range(3):
        info, iters = plan2.run(np.ones(S), -gam)
        assert np.all(iters == 0)
        Xe, _, _ = lhl_hip.shift_solve_batched(ws64, np.ones(S), -gam, Bh, offs)
        assert np.linalg.norm(X.cpu().numpy() - Xe) / np.linalg.norm(Xe) <= 1e-12 * n
    # argument checks
    with pytest.raises(lhl_hip.ArgumentError):
        lhl_hip.BatchPlan(ws64, offs, Bh, Bh)                  # host arrays
    with pytest.raises(lhl_hip.DimensionMismatch):
        plan2.run(np.ones(S + 1), -np.ones(S + 1))
    with pytest.raises(lhl_hip.ArgumentError):
        plan2.run(np.ones(S) * 1j, -np.ones(S))                # complex on a real plan
