"""Pins the oracle's quasar restatement (oracle/cpu_ref.c, src/quasar.c) - CPU only.

* bit-identical to the compiled reference (a private copy of oracle/_ref's
  library per run, so its static random stream and knot phase start from the
  reference's initial state) over 40 steps of a two-black-hole scenario with
  accretion, drag, gas conversion, feedback, jet decay with lobes, recycling,
  spawning, capacity limits and compaction;
* the reference's quasar test assertions (tests/test_physics.c:423-575, 700-836).
"""
import numpy as np
import pytest

from oracle.cpu_ref import CpuRef, QuasarConfig, RefBuild
from tests._quasar_cases import GAS, JET, LOBE, SMBH, STAR, drift, quasar_bodies
from tests._util import bodies_identical, first_body_difference

ref = CpuRef()
needs_ref = pytest.mark.skipif(not RefBuild.available(), reason="compiled reference not present")


@needs_ref
@pytest.mark.parametrize("seed,cap_extra,dt", [(1, 1500, 0.005), (2, 30, 0.02), (3, 0, 0.005)])
def test_oracle_equals_compiled_reference(seed, cap_extra, dt):
    from oracle.cpu_ref import RefQuasar
    R = RefQuasar()
    b, n = quasar_bodies(seed, cap=3000 + cap_extra)
    a = b.copy()
    ca, cb = QuasarConfig.default(jet_count=40), QuasarConfig.default(jet_count=40)
    state = [1, 0]
    na, nb = n, n
    for step in range(40):
        na = ref.quasar_step(a, na, ca, dt, state)
        nb = R.quasar_step(b, nb, cb, dt)
        assert na == nb and ca.jet_count == cb.jet_count, step
        assert bodies_identical(a[:na], b[:nb]), (step, first_body_difference(a[:na], b[:nb]))
        if step % 13 == 12:
            na = ref.quasar_compact(a, na)
            nb = R.quasar_compact(b, nb)
            assert na == nb and bodies_identical(a[:na], b[:nb])
        drift(a, na, dt)
        drift(b, nb, dt)
    assert np.sum(a["type"][:na] == JET) > 0


def one(**kw):
    b = np.zeros(kw.pop("cap", 2), dtype=__import__("cosmosim_amd")._abi.BODY_DTYPE)
    return b


def test_reference_quasar_assertions():
    from cosmosim_amd._abi import BODY_DTYPE
    # accretion conserves mass and momentum (:425-454)
    b = np.zeros(2, dtype=BODY_DTYPE)
    b[0]["mass"], b[0]["type"], b[0]["spin_z"] = 100.0, SMBH, 1.0
    b[1]["x"], b[1]["mass"], b[1]["vx"], b[1]["type"] = 0.05, 2.0, 0.1, STAR
    m0, p0 = b["mass"].sum(), np.sum(b["mass"] * b["vx"])
    ref.quasar_step(b, 2, QuasarConfig.default(), 0.005, [1, 0])
    assert abs(b["mass"].sum() - m0) <= 1e-10 and abs(np.sum(b["mass"] * b["vx"]) - p0) <= 1e-10
    assert b[1]["mass"] == 0.0
    # feedback pushes outward, radially (:456-481)
    b = np.zeros(2, dtype=BODY_DTYPE)
    b[0]["type"], b[0]["mass"], b[0]["accretion_rate"], b[0]["spin_z"] = SMBH, 100.0, 50.0, 1.0
    b[1]["x"], b[1]["mass"], b[1]["type"] = 5.0, 1.0, STAR
    ref.quasar_step(b, 2, QuasarConfig.default(), 0.005, [1, 0])
    assert b[1]["ax"] > 0 and abs(b[1]["ay"]) <= 1e-12 and abs(b[1]["az"]) <= 1e-12
    # Eddington cap (:483-501)
    b = np.zeros(1, dtype=BODY_DTYPE)
    b[0]["type"], b[0]["mass"], b[0]["accretion_rate"], b[0]["spin_z"] = SMBH, 100.0, 1e4, 1.0
    cfg = QuasarConfig.default()
    ref.quasar_step(b, 1, cfg, 0.005, [1, 0])
    assert b[0]["luminosity"] <= cfg.eddington_k * b[0]["mass"] + 1e-10
    # jets along the spin axis (:503-530)
    b = np.zeros(10, dtype=BODY_DTYPE)
    b[0]["type"], b[0]["mass"], b[0]["accretion_rate"], b[0]["spin_z"] = SMBH, 100.0, 50.0, 1.0
    n = ref.quasar_step(b, 1, QuasarConfig.default(jet_cap=10), 0.005, [1, 0])
    assert n > 1
    for i in range(1, n):
        vs = abs(b[i]["vy"]) + abs(b[i]["vz"])
        assert b[i]["type"] == JET and vs > abs(b[i]["vx"]) and vs > 1.0
    # compaction keeps live bodies in order (:553-572)
    b = np.zeros(5, dtype=BODY_DTYPE)
    b["mass"] = [1.0, 0.0, 3.0, 0.0, 5.0]
    assert ref.quasar_compact(b, 5) == 3 and list(b["mass"][:3]) == [1.0, 3.0, 5.0]
    # a distant jet is recycled as gas near the accretion radius (:702-733)
    b = np.zeros(2, dtype=BODY_DTYPE)
    b[0]["type"], b[0]["mass"], b[0]["spin_z"] = SMBH, 100.0, 1.0
    b[1]["type"], b[1]["mass"], b[1]["x"], b[1]["lifetime"] = JET, 0.5, 500.0, 1.0
    cfg = QuasarConfig.default(jet_count=1)
    ref.quasar_step(b, 2, cfg, 0.005, [1, 0])
    d = np.sqrt((b[1]["x"] - b[0]["x"]) ** 2 + (b[1]["y"] - b[0]["y"]) ** 2 +
                (b[1]["z"] - b[0]["z"]) ** 2)
    assert b[1]["type"] == GAS and d < cfg.accretion_radius * 2.0
    # ring spawning and both jet directions (:735-799)
    b = np.zeros(200, dtype=BODY_DTYPE)
    b[0]["type"], b[0]["mass"], b[0]["spin_y"], b[0]["spin_z"] = SMBH, 100.0, 0.7071, 0.7071
    b[0]["luminosity"], b[0]["accretion_rate"] = 10.0, 100.0
    n = ref.quasar_step(b, 1, QuasarConfig.default(jet_cap=200, jet_ring_count=4), 0.005, [1, 0])
    jets = b[1:n][b["type"][1:n] == JET]
    assert len(jets) >= 6
    vs = jets["vy"] * 0.7071 + jets["vz"] * 0.7071
    assert np.any(vs > 1.0) and np.any(vs < -1.0)
    # a dying jet leaves a lobe (:801-834)
    b = np.zeros(10, dtype=BODY_DTYPE)
    b[0]["type"], b[0]["mass"], b[0]["spin_z"] = SMBH, 100.0, 1.0
    b[1]["type"], b[1]["mass"], b[1]["x"], b[1]["vx"], b[1]["lifetime"] = JET, 0.5, 20.0, 5.0, 0.001
    n = ref.decay_jets(b, 2, QuasarConfig.default(jet_count=1), 0.01, [1, 0])
    assert b[1]["mass"] == 0.0 and np.sum(b["type"][:n] == LOBE) >= 1
    assert abs(b[2]["x"] - 20.0) <= 1.0
