"""PyTorch autograd harness around the library's adjoint (test harness only).

``Trajectory.apply(x0, v0, sim, substeps, every, mode)`` runs ``substeps`` KDK
substeps on the device from (x0, v0) and returns (x_k, v_k); its backward calls
``cosmosim_ext_backward`` (the hand-written adjoint: frozen trees and opening
decisions, transposed interaction lists, reverse KDK with checkpoint recompute
(mode 0) or time-reversed leapfrog (mode 1)).  Tensors are float64 on the host;
torch only carries the gradients between the loss and the C-ABI.

The merger's orbit enters through ``orbit_to_state``: body i sits at
local_i + coef_i * r_rel with velocity vlocal_i + coef_i * v_rel
(``cosmosim_ext_orbit_coefficients``), so d(state)/d(orbit) is linear.
"""
import numpy as np
import torch

import cosmosim_amd as ca


class Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, sim, substeps, every, mode):
        sim.set_state(pos=x0.detach().cpu().numpy(), vel=v0.detach().cpu().numpy())
        sim.compute_forces()            # a_0 = forces(x_0) before the first kick
        sim.record(substeps, every)
        pos, vel, _, _ = sim.state()
        ctx.sim, ctx.mode = sim, mode
        return torch.from_numpy(pos), torch.from_numpy(vel)

    @staticmethod
    def backward(ctx, gx, gv):
        gx = np.zeros((ctx.sim.n, 3)) if gx is None else gx.detach().cpu().numpy()
        gv = None if gv is None else gv.detach().cpu().numpy()
        xb, vb, _ = ctx.sim.backward(gx, gv, mode=ctx.mode)
        return torch.from_numpy(xb), torch.from_numpy(vb), None, None, None, None


def orbit_to_state(orbit, local_pos, local_vel, coef):
    """Initial state as a differentiable function of the 6 orbit parameters."""
    c = torch.as_tensor(coef, dtype=torch.float64)[:, None]
    return (torch.as_tensor(local_pos) + c * orbit[None, :3],
            torch.as_tensor(local_vel) + c * orbit[None, 3:])


def snapshot_loss(x, target):
    """Mean squared distance between final positions and a target snapshot."""
    return ((x - torch.as_tensor(target)) ** 2).sum() / x.shape[0]


def merger_problem(n, seed, orbit=None):
    """Seeded merger split into the orbit-independent part and the coefficients."""
    orbit = ca.default_orbit() if orbit is None else np.asarray(orbit, dtype=np.float64)
    b = ca.generate(ca.IC_MERGER, n, seed, orbit=orbit)
    coef = ca.orbit_coefficients(n)
    pos = ca.positions(b) - coef[:, None] * orbit[None, :3]
    vel = ca.velocities(b) - coef[:, None] * orbit[None, 3:]
    return b, pos, vel, coef, orbit
