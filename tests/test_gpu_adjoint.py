"""GPU: the hand-written adjoint against finite differences and the FP64
direct-sum adjoint.

Bars (the merger of 8,192 bodies, 50 substeps, dt 1e-3, eps 0.02, loss = mean
squared distance to a target made from orbit parameters perturbed by 1%):
  * force VJP vs central differences of the frozen-tree forces: <= 1e-6 relative;
  * tree VJP at theta = 0 vs the O(N^2) direct-sum VJP: <= 1e-12;
  * orbit-parameter gradient vs central differences (step 1e-6 relative), for the
    direct sum (theta = 0) and for the fixed-tree BH adjoint at theta = 0.5
    (differences of the frozen-tree replay): <= 1e-4 relative;
  * checkpoint recompute and time-reversed leapfrog agree to 1e-10.
"""
import numpy as np
import pytest
import torch

import cosmosim_amd as ca
from cosmosim_amd.adjoint import Trajectory, merger_problem, orbit_to_state, snapshot_loss

pytestmark = pytest.mark.gpu
N3, STEPS3, DT3, EPS3 = 8192, 50, 1e-3, 0.02


@pytest.mark.parametrize("theta", [0.0, 0.5])
def test_force_vjp_matches_frozen_finite_differences(theta):
    rng = np.random.default_rng(1)
    b = ca.generate(ca.IC_DISK, 1024, 3)
    sim = ca.Simulation(b, softening=0.02, theta=theta, precision="fp64")
    x0 = sim.state()[0]
    w = rng.normal(size=(1024, 3))
    xb = sim.force_vjp(w)
    for _ in range(3):
        d = rng.normal(size=(1024, 3))
        h = 1e-6
        fd = (np.sum(w * sim.forces_frozen(x0 + h * d))
              - np.sum(w * sim.forces_frozen(x0 - h * d))) / (2 * h)
        assert abs(fd - np.sum(xb * d)) <= 1e-6 * abs(fd)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tree_vjp_equals_direct_vjp_at_theta0(precision):
    """At theta = 0 every interaction is body-body; the tape of either walk
    (FP64 or the FP32 wave-mask walk) then gives the direct-sum adjoint."""
    rng = np.random.default_rng(2)
    b = ca.generate(ca.IC_MERGER, 2048, 5)
    b["mass"][::97] = 0.0
    sim = ca.Simulation(b, softening=0.02, theta=0.0, precision=precision)
    w = rng.normal(size=(2048, 3))
    xt = sim.force_vjp(w)
    xd = sim.direct_vjp(w)
    assert np.max(np.abs(xt - xd)) <= 1e-12 * np.max(np.abs(xd))
    assert np.all(xt[b["mass"] <= 0] == 0.0)


def _config3(theta, keep_trees):
    b, lpos, lvel, coef, orbit = merger_problem(N3, 4)
    target = ca.positions(ca.generate(ca.IC_MERGER, N3, 4, orbit=orbit * 1.01))
    sim = ca.Simulation(b, softening=EPS3, theta=theta, dt=DT3, precision="fp64")
    return sim, lpos, lvel, coef, orbit, target


def _loss_of(sim, orbit, lpos, lvel, coef, target, theta, replay):
    x0, v0 = orbit_to_state(torch.as_tensor(orbit), lpos, lvel, coef)
    if replay:
        p, _ = sim.replay(x0.numpy(), v0.numpy())
    else:
        sim.set_state(pos=x0.numpy(), vel=v0.numpy())
        sim.compute_forces()
        sim.advance(STEPS3)
        p = sim.state()[0]
    return float(np.sum((p - target) ** 2) / N3)


@pytest.mark.parametrize("theta", [0.0, 0.5])
def test_orbit_gradient_matches_finite_differences(theta):
    sim, lpos, lvel, coef, orbit, target = _config3(theta, theta > 0)
    orb = torch.tensor(orbit, requires_grad=True)
    x0, v0 = orbit_to_state(orb, lpos, lvel, coef)
    if theta > 0:   # keep every tree so the differences can replay them frozen
        sim.set_state(pos=x0.detach().numpy(), vel=v0.detach().numpy())
        sim.compute_forces()
        sim.record(STEPS3, 10, keep_trees=True)
        xk = torch.from_numpy(sim.state()[0])
        loss = snapshot_loss(xk, target)
        xb, vb, _ = sim.backward(2 * (xk.numpy() - target) / N3, mode=0)
        x0.backward(torch.from_numpy(xb), retain_graph=True)
        v0.backward(torch.from_numpy(vb))
    else:
        xk, vk = Trajectory.apply(x0, v0, sim, STEPS3, 10, 0)
        loss = snapshot_loss(xk, target)
        loss.backward()
    grad = orb.grad.numpy()
    fd = np.zeros(6)
    for k in range(6):
        h = 1e-6 * max(1.0, abs(orbit[k]))
        lp, lm = orbit.copy(), orbit.copy()
        lp[k] += h
        lm[k] -= h
        fd[k] = (_loss_of(sim, lp, lpos, lvel, coef, target, theta, theta > 0)
                 - _loss_of(sim, lm, lpos, lvel, coef, target, theta, theta > 0)) / (2 * h)
    rel = np.abs(grad - fd) / np.abs(fd)
    assert np.all(rel <= 1e-4), (grad, fd, rel)
    assert float(loss.detach()) > 0


def test_initial_state_gradient_matches_finite_differences():
    sim, lpos, lvel, coef, orbit, target = _config3(0.0, False)
    x0, v0 = orbit_to_state(torch.as_tensor(orbit), lpos, lvel, coef)
    x0 = x0.clone().requires_grad_(True)
    v0 = v0.clone().requires_grad_(True)
    xk, vk = Trajectory.apply(x0, v0, sim, STEPS3, 25, 0)
    snapshot_loss(xk, target).backward()
    gx, gv = x0.grad.numpy(), v0.grad.numpy()
    rng = np.random.default_rng(3)
    for body in rng.choice(N3, 3, replace=False):
        for arr, g in ((x0, gx), (v0, gv)):
            base = arr.detach().numpy().copy()
            h = 1e-6
            vals = []
            for sgn in (1, -1):
                p = base.copy()
                p[body, 0] += sgn * h
                xx = p if arr is x0 else x0.detach().numpy()
                vv = p if arr is v0 else v0.detach().numpy()
                sim.set_state(pos=xx, vel=vv)
                sim.compute_forces()
                sim.advance(STEPS3)
                vals.append(np.sum((sim.state()[0] - target) ** 2) / N3)
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[body, 0]) <= 1e-4 * abs(fd) + 1e-12


@pytest.mark.parametrize("every", [10, 25])
def test_checkpoint_recompute_and_time_reversal_agree(every):
    sim, lpos, lvel, coef, orbit, target = _config3(0.5, False)
    sim.record(STEPS3, every)
    xk = sim.state()[0]
    g = 2 * (xk - target) / N3
    x0a, v0a, st0 = sim.backward(g, mode=0)
    x0b, v0b, st1 = sim.backward(g, mode=1)
    assert np.max(np.abs(x0a - x0b)) <= 1e-10 * np.max(np.abs(x0a))
    assert np.max(np.abs(v0a - v0b)) <= 1e-10 * np.max(np.abs(v0a))
    assert st1[0] < 1e-12          # reversed positions vs the checkpoints
    assert st0[1] > 0
    # the handle is left at the recorded final state
    assert np.array_equal(sim.state()[0], xk)
