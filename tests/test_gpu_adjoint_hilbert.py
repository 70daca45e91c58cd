"""GPU: adjoint tapes recorded with Peano-Hilbert wavefronts.

A taped FP32 walk groups bodies into wavefronts as the untaped walk does (Peano-
Hilbert order when the tree has it and there are more wavefronts than the GPU
holds at once), so its forces are the trajectory's bit for bit; the tape's lanes
then map back to sorted bodies through the build's rank-to-position table.  The
environment lowers the thresholds so that these small inputs take that path.
Checks:
  * the force VJP of a Peano-Hilbert tape equals that of a sorted-order tape (the
    same interactions, summed in another order): <= 1e-13 relative; at theta = 0
    it equals the direct-sum VJP (<= 1e-12), with dead bodies;
  * the frozen-tape forces do not depend on the grouping (bit-identical: each body
    sums its own interactions in its own visiting order);
  * a taped walk's accelerations equal the untaped walk's bit for bit;
  * checkpoint recompute and time reversal agree (<= 1e-10), the reversal stays
    within 1e-12 of the checkpoints, and the gradient agrees with the sorted-order
    run to the FP32 trajectory's rounding (<= 1e-6);
  * the frozen-tree replay of a recording with Peano-Hilbert tapes equals the one
    with sorted-order tapes (<= 1e-12).
"""
import os

import numpy as np
import pytest

import cosmosim_amd as ca

pytestmark = pytest.mark.gpu

ENV = ("COSMOSIM_HILBERT_WAVES", "COSMOSIM_HILBERT_MIN_BODIES", "COSMOSIM_HILBERT_MIN_WAVES")


def _sim(b, hilbert, **kw):
    """A handle whose FP32 walks (taped ones included) use Peano-Hilbert wavefronts
    for any input (hilbert) or sorted-order wavefronts."""
    env = dict(zip(ENV, ("1" if hilbert else "0", "1", "0")))
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return ca.Simulation(b, **kw)
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def _merger(n, seed, dead=0):
    b = ca.generate(ca.IC_MERGER, n, seed)
    if dead:
        b["mass"][::dead] = 0.0
    return b


def _uses_hilbert(sim):
    si, wi, _ = sim.debug_orders()
    return not np.array_equal(si, wi)


@pytest.mark.parametrize("theta", [0.0, 0.5])
def test_vjp_of_hilbert_tape(theta):
    n = 6000 if theta == 0.0 else 20000
    b = _merger(n, 7, dead=97)
    rng = np.random.default_rng(4)
    w = rng.normal(size=(n, 3))
    kw = dict(softening=0.01, theta=theta, precision="fp32")
    s1, s0 = _sim(b, True, **kw), _sim(b, False, **kw)
    x1, x0 = s1.force_vjp(w), s0.force_vjp(w)
    assert _uses_hilbert(s1) and not _uses_hilbert(s0)
    scale = np.max(np.abs(x0))
    print(f"\ntheta {theta}: Hilbert vs sorted tape VJP {np.max(np.abs(x1 - x0)) / scale:.2e}")
    assert np.max(np.abs(x1 - x0)) <= 1e-13 * scale
    assert np.all(x1[b["mass"] <= 0] == 0.0)
    if theta == 0.0:
        xd = s1.direct_vjp(w)
        assert np.max(np.abs(x1 - xd)) <= 1e-12 * np.max(np.abs(xd))
    s1.close()
    s0.close()


def test_frozen_forces_of_hilbert_tape_are_grouping_free():
    b = _merger(20000, 8, dead=89)
    rng = np.random.default_rng(5)
    w = rng.normal(size=(20000, 3))
    pos = ca.positions(b) + rng.normal(size=(20000, 3)) * 1e-6
    kw = dict(softening=0.01, theta=0.6, precision="fp32")
    out = []
    for h in (True, False):
        s = _sim(b, h, **kw)
        s.force_vjp(w)                        # build + taped walk at b's positions
        assert _uses_hilbert(s) == h
        out.append(s.forces_frozen(pos))
        s.close()
    assert np.array_equal(out[0].view(np.uint64), out[1].view(np.uint64))


def test_taped_walk_forces_equal_untaped():
    """The recording's taped walks give the accelerations the untaped substeps
    give: a recorded run and a plain run of the same substeps end bit-identical."""
    b = _merger(20000, 9, dead=101)
    kw = dict(softening=0.01, theta=0.5, dt=5e-4, precision="fp32")
    s1 = _sim(b, True, **kw)
    s1.compute_forces()
    s1.record(6, 3, keep_trees=True)          # taped walks throughout
    s2 = _sim(b, True, **kw)
    s2.compute_forces()
    s2.advance(6)                             # untaped walks
    for a, c in zip(s1.state(), s2.state()):
        assert np.array_equal(np.asarray(a).view(np.uint64), np.asarray(c).view(np.uint64))
    s1.close()
    s2.close()


def test_time_reversal_with_hilbert_tapes():
    n, steps = 20000, 20
    b = _merger(n, 10, dead=113)
    rng = np.random.default_rng(6)
    g = rng.normal(size=(n, 3)) / n
    kw = dict(softening=0.01, theta=0.5, dt=5e-4, precision="fp32")
    grads = []
    for h in (True, False):
        s = _sim(b, h, **kw)
        s.compute_forces()
        s.record(steps, 10)
        xa, va, st0 = s.backward(g, mode=0)
        xb, vb, st1 = s.backward(g, mode=1)
        assert np.max(np.abs(xa - xb)) <= 1e-10 * np.max(np.abs(xa))
        assert np.max(np.abs(va - vb)) <= 1e-10 * np.max(np.abs(va))
        assert st1[0] < 1e-12
        grads.append((xa, va))
        s.close()
    for k in range(2):
        ref = grads[1][k]
        rel = np.max(np.abs(grads[0][k] - ref)) / np.max(np.abs(ref))
        print(f"\nHilbert vs sorted tapes, gradient {k}: {rel:.2e}")
        assert rel <= 1e-6


def test_replay_of_hilbert_tapes():
    n = 20000
    b = _merger(n, 11, dead=71)
    kw = dict(softening=0.01, theta=0.5, dt=5e-4, precision="fp32")
    out = []
    for h in (True, False):
        s = _sim(b, h, **kw)
        s.compute_forces()
        x0, v0 = s.state()[0].copy(), s.state()[1].copy()
        s.record(2, 1, keep_trees=True)
        p, v = s.replay(x0, v0)
        out.append((p, v))
        s.close()
    for k in range(2):
        ref = out[1][k]
        assert np.max(np.abs(out[0][k] - ref)) <= 1e-12 * np.max(np.abs(ref))
