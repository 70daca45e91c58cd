"""GPU: the eight exports of src/cosmosim_api.h through a library built with
initial-condition generators (tests/_lib/libcosmosim_ic_stub.so: the same core
linked with the test-only generators of tests/ic_stub.c, as a caller links its
own src/initial_conditions.c with the build's --ic-source option).

  * tests/test_physics.c:1048-1137 restated: handles in every mode, the
    quasar-mode assertions of test_api_quasar_mode (an SMBH exists, five steps,
    bodies stay active);
  * cosmosim_step in non-quasar mode against the oracle's KDK substeps from
    the same bodies (G 1, softening 1.5): positions within 1e-12;
  * cosmosim_step in quasar mode (the plugin's MergerMode + QuasarMode), 60
    calls of two substeps each, compaction at the 60th call
    (src/cosmosim_api.c:115-144): before each call the oracle takes the
    device's bodies, random state and quasar config and composes the
    SPH-enabled integrator step (src/integrator.c:12-57 with src/sph.c) and
    quasar_step (src/quasar.c:477-510) per substep, then quasar_compact; body
    count, types, random state and jet count identical, every float field
    within 1e-9 (theta 0, where the gravity walk is bit-exact; jets spawned in
    the first substep carry sin/cos/log rounding into the second);
  * the default library, built without generators, refuses cosmosim_create
    with COSMOSIM_E_UNSUPPORTED instead of generating other bodies.
"""
import os

import numpy as np
import pytest

import cosmosim_amd as ca
from oracle.cpu_ref import CpuRef, QuasarConfig as RefConfig
from tests._util import BODY_FLOAT_FIELDS, acc, rel_err

pytestmark = pytest.mark.gpu
ref = CpuRef()
STUB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib",
                    "libcosmosim_ic_stub.so")
SMBH = 2
G, EPS = 1.0, 1.5                     # fixed inside the handle API (src/cosmosim_api.c:10-11)


def _pos(b):
    return np.stack([b["x"], b["y"], b["z"]], 1)


def test_api_handle_like_reference_tests():
    """tests/test_physics.c:1048-1111 restated."""
    h = ca.Handle(library=STUB, n_bodies=1000, merger=0, quasar=0)
    assert h.count() == 1000 and h.active_count() == 1000 and h.sim_time() == 0.0
    b0 = h.bodies()
    h.step()
    assert abs(h.sim_time() - h.config.dt * h.config.substeps) < 1e-12
    b1 = h.bodies()
    assert np.any(b1["vx"] != b0["vx"])
    h.close()
    h = ca.Handle(library=STUB, n_bodies=1, merger=1)   # n < 2 becomes 2 (:49-51)
    assert h.count() == 2
    h.close()


def test_api_quasar_mode_like_reference_test():
    """tests/test_physics.c:1105-1137 restated (the plugin's default modes)."""
    h = ca.Handle(library=STUB, n_bodies=1000, merger=1, quasar=1)
    b = h.bodies()
    assert h.count() == 1000 + 200                       # n + n/5 dust
    assert np.any((b["type"] == SMBH) & (b["mass"] > 0))
    for _ in range(5):
        h.step()
    assert h.active_count() > 0
    assert np.isfinite(_pos(h.bodies())).all()
    h.close()


def test_api_step_matches_oracle_substeps():
    h = ca.Handle(library=STUB, n_bodies=2000, merger=1, quasar=0, substeps=3)
    start = h.bodies()
    h.step()
    got = h.bodies()
    bb = start.copy()
    for _ in range(3):
        ref.step(bb, h.config.dt, G, EPS, h.config.theta)
    assert np.max(np.abs(_pos(got) - _pos(bb))) <= 1e-12 * max(1.0, np.max(np.abs(_pos(bb))))
    assert rel_err(acc(got), acc(bb)).max() <= 1e-10
    h.close()


def _ref_config(qc):
    rc = RefConfig()
    for f, _ in RefConfig._fields_:
        setattr(rc, f, getattr(qc, f))
    return rc


def _oracle_sph_substep(w, n, dt, theta):
    """integrator_step with SPH (src/integrator.c:12-57) on w[:n]."""
    v = w[:n]
    ref.kick_drift(v, dt)
    pool = ref.build(v)
    ref.forces(pool, v, G, EPS, theta)
    ref.sph_density(v, pool)
    ref.sph_forces(v, pool)
    ref.kick(v, dt)
    ref.sph_cooling(v, dt)


def test_api_quasar_mode_matches_oracle_composition():
    h = ca.Handle(library=STUB, n_bodies=1000, merger=1, quasar=1, theta=0.0)
    n_alloc = 1000 + 200 + 250                            # n + n/5 + n/4 (:53-54)
    dt, theta = h.config.dt, h.config.theta
    for call in range(60):
        start = h.bodies()
        n = len(start)
        w = np.zeros(n_alloc, dtype=ca.BODY_DTYPE)
        w[:n] = start
        state = h.quasar_state()
        rcfg = _ref_config(h.quasar_config())
        assert rcfg.max_bodies == n_alloc and rcfg.jet_cap == 100
        for _ in range(h.config.substeps):
            _oracle_sph_substep(w, n, dt, theta)
            n = ref.quasar_step(w, n, rcfg, dt, state)
        if call == 59:
            n = ref.quasar_compact(w, n)
        h.step()
        got = h.bodies()
        assert h.count() == n, call
        assert h.quasar_state() == state, call
        assert h.quasar_config().jet_count == rcfg.jet_count, call
        want = w[:n]
        assert np.array_equal(got["type"], want["type"]), call
        for f in BODY_FLOAT_FIELDS:
            d = np.abs(got[f] - want[f])
            assert np.all(d <= 1e-9 * np.maximum(1.0, np.abs(want[f]))), (call, f)
    assert abs(h.sim_time() - 60 * dt * h.config.substeps) < 1e-12
    assert h.active_count() > 0
    h.close()


def test_default_library_without_generators_refuses_create():
    with pytest.raises(ca.CosmosimError) as e:
        ca.Handle(n_bodies=1000, merger=1, quasar=1)
    assert e.value.code == -7
    with pytest.raises(ca.CosmosimError) as e:
        ca.Handle(n_bodies=1000, merger=0, quasar=0)
    assert e.value.code == -7
