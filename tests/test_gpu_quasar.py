"""GPU: quasar physics on the device (cosmosim_ext_quasar_step / _compact)
against the oracle's restatement of src/quasar.c (pinned to the compiled
reference by tests/test_oracle_quasar.py).

Bars:
  * one step from the same state (bodies, config, random state, knot phase):
    body count, jet count, random state and every field of every body
    bit-identical - except the positions and velocities (and, for recycled
    jets, the internal energy derived from them) of spawned and recycled jets,
    which go through sin/cos/log: those within 1e-13 relative;
  * 25 device-resident steps (no resync) with compaction: counts, types and
    masses identical, positions within 1e-9;
  * SPH substeps plus quasar steps through the extension API against the
    oracle's composition of the same passes: within 1e-9.
"""
import numpy as np
import pytest

import cosmosim_amd as ca
from oracle.cpu_ref import CpuRef, QuasarConfig as RefConfig
from tests._quasar_cases import GAS, JET, drift, quasar_bodies
from tests._util import BODY_FLOAT_FIELDS, same_bits

pytestmark = pytest.mark.gpu
ref = CpuRef()
VEC = ["x", "y", "z", "vx", "vy", "vz"]


def cfg_pair(**kw):
    return ca.QuasarConfig.default(**kw), RefConfig.default(**kw)


def device_one_step(b, n, cfg, dt, state):
    sim = ca.Simulation(b[:n], softening=0.01, theta=0.5, precision="fp64")
    sim.reserve(len(b))
    sim.set_state(acc=np.stack([b["ax"][:n], b["ay"][:n], b["az"][:n]], 1))
    sim.quasar_state(*state)
    n2 = sim.quasar_step(cfg, dt)
    out = sim.bodies()
    st = sim.quasar_state()
    sim.close()
    return out, n2, st


def assert_step_equal(got, want, before, n_before):
    """Bit-identical except the transcendental fields of spawned / recycled jets."""
    assert len(got) == len(want)
    assert np.array_equal(got["type"], want["type"])
    loose = np.zeros(len(want), bool)
    loose[n_before:] = want["type"][n_before:] == JET                      # spawned jets
    m = min(len(before), len(want))
    recycled = np.zeros(len(want), bool)
    recycled[:m] = (before["type"][:m] == JET) & (want["type"][:m] == GAS)
    loose |= recycled
    for f in BODY_FLOAT_FIELDS:
        tight = ~loose if f in VEC else ~recycled if f == "internal_energy" else np.ones_like(loose)
        assert same_bits(got[f][tight], want[f][tight]), (f, np.nonzero(
            got[f][tight].view(np.uint64) != want[f][tight].view(np.uint64))[0][:5])
    for f in VEC + ["internal_energy"]:
        sel = loose if f in VEC else recycled
        if sel.any():
            d = np.abs(got[f][sel] - want[f][sel])
            assert np.all(d <= 1e-13 * np.maximum(1.0, np.abs(want[f][sel]))), f


@pytest.mark.parametrize("seed,extra,dt", [(1, 1500, 0.005), (2, 30, 0.02), (3, 0, 0.005)])
def test_each_step_equals_oracle(seed, extra, dt):
    b, n = quasar_bodies(seed, cap=3000 + extra)
    ref_cfg = RefConfig.default(jet_count=40)
    state = [1, 0]
    for step in range(12):
        dcfg = ca.QuasarConfig.default(jet_count=ref_cfg.jet_count)
        before = b[:n].copy()
        got, n_dev, st_dev = device_one_step(b, n, dcfg, dt, state)
        n2 = ref.quasar_step(b, n, ref_cfg, dt, state)
        assert n_dev == n2 and dcfg.jet_count == ref_cfg.jet_count, step
        assert st_dev == state, step
        assert_step_equal(got, b[:n2], before, n)
        n = n2
        drift(b, n, dt)


def test_device_resident_steps_with_compaction():
    b, n = quasar_bodies(4, cap=4000)
    sim = ca.Simulation(b[:n], softening=0.01, theta=0.5, precision="fp64")
    sim.reserve(len(b))
    sim.set_state(acc=np.stack([b["ax"][:n], b["ay"][:n], b["az"][:n]], 1))
    dcfg, rcfg = cfg_pair(jet_count=40)
    state = [1, 0]
    for step in range(25):
        nd = sim.quasar_step(dcfg, 0.005)
        n = ref.quasar_step(b, n, rcfg, 0.005, state)
        assert nd == n and dcfg.jet_count == rcfg.jet_count, step
        if step % 10 == 9:
            assert sim.quasar_compact() == ref.quasar_compact(b, n)
            n = sim.count()
    got = sim.bodies()
    assert sim.quasar_state() == state
    want = b[:n]
    assert np.array_equal(got["type"], want["type"]) and same_bits(got["mass"], want["mass"])
    for f in VEC:
        assert np.all(np.abs(got[f] - want[f]) <= 1e-9 * np.maximum(1.0, np.abs(want[f]))), f
    sim.close()


def test_sph_and_quasar_substeps_through_the_api():
    b, n = quasar_bodies(5, n=2500, cap=3500)
    b["mass"][2:n] = np.where(b["type"][2:n] == JET, b["mass"][2:n], b["mass"][2:n] * 0.01)
    sim = ca.Simulation(b[:n], softening=0.05, theta=0.0, dt=1e-3, precision="fp64_strict")
    sim.reserve(len(b))
    w = b.copy()
    acc = sim.accelerations()
    w["ax"][:n], w["ay"][:n], w["az"][:n] = acc.T
    dcfg, rcfg = cfg_pair(jet_count=40)
    state = [1, 0]
    for _ in range(3):
        sim.advance_sph(1)
        nd = sim.quasar_step(dcfg, 1e-3)
        ref.kick_drift(w[:n], 1e-3)
        pool = ref.build(w[:n])
        ref.forces(pool, w[:n], 1.0, 0.05, 0.0)
        ref.sph_density(w[:n], pool)
        ref.sph_forces(w[:n], pool)
        ref.kick(w[:n], 1e-3)
        ref.sph_cooling(w[:n], 1e-3)
        n = ref.quasar_step(w, n, rcfg, 1e-3, state)
        assert nd == n and dcfg.jet_count == rcfg.jet_count
    got = sim.bodies()
    assert np.array_equal(got["type"], w["type"][:n])
    for f in VEC + ["density", "smoothing_h", "internal_energy", "luminosity"]:
        d = np.abs(got[f] - w[f][:n])
        assert np.all(d <= 1e-9 * np.maximum(1.0, np.abs(w[f][:n]))), f
    sim.close()
