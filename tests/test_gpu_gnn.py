"""Learned warm start on the GPU: the device network against the CPU
restatement (oracle/gnn_ref.py), training, determinism and the fallback.

Tolerances:
  * loss: 1e-4 relative; parameter gradient: per parameter block,
    max |g_dev - g_ref| <= 5e-3 max |g_ref|.  The oracle rounds the operands
    of the device's BF16 MFMA products exactly as the device does, so what
    remains is FP32 versus FP64 accumulation and the rare BF16 rounding or
    ReLU decision that flips between the two;
  * training: bit-identical parameters from the same seed; the loss falls;
  * the prediction is the previous update plus the network's correction: an
    all-zero network guesses zero on the first step (bit-identical to the
    zero-guess path, the solve keeps it: its initial residual equals ||b||)
    and the previous update afterwards (1e-12 relative).
"""
import numpy as np
import pytest

import cfdg
import cfdg.cases as cases
import gnn_ref
from helpers import perturbed_state

pytestmark = pytest.mark.gpu


def _recorded(steps=2, seed=1, ni=24, nj=16, nk=8, gnn_seed=3):
    case = cases.plate(ni=ni, nj=nj, nk=nk, seed=seed)
    mesh = cfdg.Mesh.generate(**case.mesh)
    s = cfdg.Solver(mesh, case.numerics, case.bcs)
    U = perturbed_state(s, case, seed=seed)  # init_uniform sets the freestream
    s.set_state(U)
    s.set_recording(True)
    g = cfdg.Gnn(seed=gnn_seed)
    lp = cases.default_linparams()
    for _ in range(steps):
        s.implicit_step(lp)
        g.record(s)
    return case, mesh, s, g


def _blocks():
    o = 0
    for name, shape in gnn_ref.SIZES:
        n = int(np.prod(shape))
        yield name, slice(o, o + n)
        o += n


def test_loss_and_gradient_match_cpu_oracle():
    case, mesh, s, g = _recorded(steps=2)
    assert g.num_samples() == 2
    p = g.params().astype(np.float64)
    assert np.all(p.reshape(-1)[: 128 * 32].reshape(128, 32)[:, gnn_ref.FUSED:] == 0.0)
    for k in range(2):
        loss, grad = g.loss_grad(k)
        smp = g.sample(k)
        assert np.all(smp["x"][:, gnn_ref.FUSED:] == 0.0)
        ref_loss, ref_grad = gnn_ref.loss_grad(p, smp["x"], smp["t"], smp["adj_off"], smp["adj"], smp["rev"],
                                               smp["eg"].astype(np.float64), quant=True)
        assert abs(loss - ref_loss) <= 1e-4 * abs(ref_loss), (loss, ref_loss)
        for name, sl in _blocks():
            d = np.abs(grad[sl] - ref_grad[sl]).max()
            assert d <= 5e-3 * np.abs(ref_grad[sl]).max() + 1e-12, (name, d, np.abs(ref_grad[sl]).max())


def test_graph_is_symmetric_with_boundary_self_loops():
    case, mesh, s, g = _recorded(steps=1)
    smp = g.sample(0)
    off, adj, rev, eg = smp["adj_off"], smp["adj"], smp["rev"], smp["eg"]
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    assert np.array_equal(adj[rev], src)          # rev is the reverse edge
    assert np.array_equal(rev[rev], np.arange(len(adj)))
    inner = adj != src
    assert np.allclose(eg[rev][inner, :3], -eg[inner, :3])  # normals flip across a face
    assert np.sum(~inner) == mesh.n_faces - mesh.n_inner_faces  # one self loop per boundary face


def test_training_lowers_loss_and_is_deterministic():
    case, mesh, s, g = _recorded(steps=4)
    p0 = g.params()
    l0 = np.mean([g.loss_grad(k)[0] for k in range(4)])
    g.train(epochs=40, lr=2e-3, seed=7)
    p1 = g.params()
    l1 = np.mean([g.loss_grad(k)[0] for k in range(4)])
    assert l1 < 0.7 * l0, (l0, l1)
    g.set_params(p0)
    g.train(epochs=40, lr=2e-3, seed=7)
    assert np.array_equal(g.params(), p1)


def test_zero_network_predicts_previous_update():
    """The network predicts the change of the update over a skip connection:
    with all parameters zero the first step's guess is zero (bit-identical to
    the zero-guess path) and later guesses are the previous step's update."""
    case = cases.plate(ni=24, nj=16, nk=8)
    mesh = cfdg.Mesh.generate(**case.mesh)
    a = cfdg.Solver(mesh, case.numerics, case.bcs)
    b = cfdg.Solver(mesh, case.numerics, case.bcs)
    U0 = perturbed_state(a, case)
    b.init_uniform(case.rho, case.velocity, case.p)
    g = cfdg.Gnn(seed=1)
    g.set_params(np.zeros(g.num_params, np.float32))
    b.set_warm_start(g)
    lp = cases.default_linparams()
    for s in (a, b):
        s.set_state(U0)
        s.implicit_step(lp)
    U1 = b.get_state()
    assert np.array_equal(a.get_state(), U1)
    assert b.warm_start_stats() == (1, 0)
    b.eval_residual()
    b.compute_dt()
    du = g.predict(b)
    ref = (U1 - U0).T.ravel()  # AoS, mesh order
    assert np.max(np.abs(du - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_learned_warm_start_effect_on_a_steady_run():
    """What the trained network does to a steady run, measured against the
    same run without it (plate 24x16x8, unlimited MUSCL, CFL 1000, to a
    relative residual of 1e-11): the run with the network converges to the
    same steady state (1e-8 of the freestream scales), every step either kept
    the prediction or fell back to the zero guess,
    and the linear iterations of both runs are reported.  The network is
    trained on two other seeded plates, every 50th of 1000 steps.

    On this problem the learned guess does not lower the iteration count:
    each solve stops at 10 % of ||b||, a zero guess removes the easily reduced
    components first, and a guess leaves the slowly reduced ones (see the
    printed counts); the test pins correctness and records the effect."""
    kw = dict(limiter="UNLIMITED", cfl=1000.0)
    lp = cases.default_linparams()

    def solver(case, seed):
        mesh = cfdg.Mesh.generate(**case.mesh)
        s = cfdg.Solver(mesh, case.numerics, case.bcs)
        s.set_state(perturbed_state(s, case, seed=seed))
        return mesh, s

    g = cfdg.Gnn(seed=11)
    for m in range(2):
        case = cases.plate(ni=24, nj=16, nk=8, mach=0.25 + 0.1 * m, seed=40 + m, **kw)
        _, s = solver(case, m + 5)
        s.set_recording(True)
        for it in range(1000):
            s.implicit_step(lp)
            if it % 50 == 49:
                g.record(s)
    g.train(epochs=30, lr=1e-3, seed=5)
    case = cases.plate(ni=24, nj=16, nk=8, **kw)
    _, plain = solver(case, 1)
    rp = plain.run_steady(lp, 12000, rel_tol=1e-11)
    _, warm = solver(case, 1)
    warm.set_warm_start(g)
    rw = warm.run_steady(lp, 12000, rel_tol=1e-11)
    assert rp.status == rw.status == cfdg.CONVERGED
    used, rejected = warm.warm_start_stats()
    assert used + rejected == rw.steps
    speed = float(np.linalg.norm(case.velocity))
    scale = np.array([case.rho, speed, speed, speed, case.p])[:, None]

    def prim(U):
        rho = U[0]
        u, v, w = U[1] / rho, U[2] / rho, U[3] / rho
        return np.stack([rho, u, v, w, (cases.GAMMA - 1.0) * (U[4] - 0.5 * rho * (u * u + v * v + w * w))])

    err = np.max(np.abs(prim(warm.get_state()) - prim(plain.get_state())) / scale)
    print(f"plain: {rp.steps} steps, {rp.linear_iterations} linear iterations; learned: {rw.steps} steps, "
          f"{rw.linear_iterations} linear iterations, kept {used}, rejected {rejected}; state difference {err:.2e}")
    assert err <= 1e-8
