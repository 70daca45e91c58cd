"""Pins the oracle's float snapshots (oracle/cpu_ref.c: the Unreal plugin's
FCosmosimBodyGPU, CosmosimSubsystem.cpp:12-43, and the renderer's upload
record, src/renderer.c:348-364) - CPU only.

Neither consumer compiles here (Unreal, OpenGL), so the restatement is pinned by
an independent construction: numpy's IEEE round-to-nearest float32 conversion
of the live rows, field by field, and the cap semantics of the plugin's loop.
"""
import numpy as np

from cosmosim_amd import generate, IC_MERGER
from oracle.cpu_ref import CpuRef

ref = CpuRef()
UE = ["x", "y", "z", "mass", "vx", "vy", "vz", "type", "internal_energy", "density", "lifetime",
      "luminosity", "spin_x", "spin_y", "spin_z"]


def bodies(seed=3, n=5000):
    rng = np.random.default_rng(seed)
    b = generate(IC_MERGER, n, seed)
    for f in ("internal_energy", "density", "lifetime", "luminosity", "spin_x", "spin_y", "spin_z"):
        b[f] = rng.normal(size=n) * 10.0 ** rng.uniform(-40, 40, n)   # far outside float range too
    b["type"] = rng.integers(0, 6, n)
    b["mass"][::7] = 0.0
    b["mass"][::11] = -1.0
    b["x"][5] = 1e300                     # overflows to inf in float
    return b


def numpy_records(b, fields):
    live = ~(b["mass"] <= 0)
    with np.errstate(over="ignore"):   # values beyond float range become inf, as in C
        cols = [np.asarray(b[f][live]).astype(np.float64).astype(np.float32) for f in fields]
    return np.stack(cols, 1)


def test_ue_records_equal_numpy_conversion():
    b = bodies()
    got = ref.snapshot(b, 0)
    want = numpy_records(b, UE)
    assert got.shape == (want.shape[0], 16)
    assert np.array_equal(got[:, :15].view(np.uint32), want.view(np.uint32))
    assert np.all(got[:, 15] == 0.0)


def test_render_records_equal_numpy_conversion():
    b = bodies(4)
    got = ref.snapshot(b, 1)
    assert np.array_equal(got.view(np.uint32), numpy_records(b, UE[:10]).view(np.uint32))


def test_ue_cap_keeps_the_first_live_bodies():
    b = bodies(5, 300)
    full = ref.snapshot(b, 0)
    for cap in [0, 1, 57, len(full), len(full) + 10]:
        got = ref.snapshot(b, 0, cap)
        assert np.array_equal(got.view(np.uint32), full[:cap].view(np.uint32))
