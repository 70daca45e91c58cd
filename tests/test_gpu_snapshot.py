"""GPU: the float snapshots built on the device (cosmosim_ext_snapshot_ue,
cosmosim_ext_snapshot_render) against the oracle's restatement of the host
consumers' loops (CosmosimSubsystem.cpp:12-43, src/renderer.c:348-364).
Bar: bit-identical records and counts, cap included, after device-resident
substeps and on an empty and an all-dead handle."""
import numpy as np
import pytest

import cosmosim_amd as ca
from oracle.cpu_ref import CpuRef

pytestmark = pytest.mark.gpu
ref = CpuRef()


def mixed(n, seed):
    rng = np.random.default_rng(seed)
    b = ca.generate(ca.IC_MERGER, n, seed)
    for f in ("internal_energy", "density", "lifetime", "luminosity", "spin_x", "spin_y", "spin_z"):
        b[f] = rng.normal(size=n) * 10.0 ** rng.uniform(-40, 40, n)
    b["type"] = rng.integers(0, 6, n)
    b["mass"][::7] = 0.0
    return b


@pytest.mark.parametrize("n", [4, 2047, 2048, 2049, 100_003])
def test_snapshots_equal_oracle(n):
    b = mixed(n, 3)
    sim = ca.Simulation(b, softening=0.01, theta=0.5, dt=1e-3, precision="fp32")
    sim.advance(2)
    now = sim.bodies()
    for got, want in ((sim.snapshot_ue(), ref.snapshot(now, 0)),
                      (sim.snapshot_render(), ref.snapshot(now, 1))):
        assert got.shape == want.shape
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    full = len(ref.snapshot(now, 0))
    for cap in [0, 1, full // 3, full + 5]:
        got = sim.snapshot_ue(cap)
        assert np.array_equal(got.view(np.uint32), ref.snapshot(now, 0, cap).view(np.uint32))
    sim.close()


def test_all_dead_and_million():
    b = ca.make_bodies(np.zeros((4, 3)), mass=[0.0, -1.0, 0.0, 0.0])
    sim = ca.Simulation(b, softening=0.01)
    assert sim.snapshot_ue().shape == (0, 16) and sim.snapshot_render().shape == (0, 10)
    sim.close()
    b = ca.generate(ca.IC_MERGER, 1_000_000, 3)
    sim = ca.Simulation(b, softening=0.005, theta=0.5, dt=5e-4, precision="fp32")
    sim.advance(1)
    got = sim.snapshot_render()
    want = ref.snapshot(sim.bodies(), 1)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    sim.close()
