"""GPU: the FP32 walk on wave-shared interaction lists (precision "fp32_list").

Each wavefront takes 64 bodies consecutive in the wavefront order, lists every
node that passes the opening test at the group's bounding box (and every leaf),
and sums the list for each body.  Checks:

  * against the oracle's restatement (cref_list_walk) on the device's own tree
    and wavefront order: per-body list counts equal, accelerations within the
    FP32 tolerance stated below; the oracle itself checks that every group's
    list covers the live mass exactly once and that every listed cell passes
    the reference criterion (2h)^2/d^2 < theta^2 for every body of its group;
  * theta = 0 lists every body: the FP64 direct sum to FP32 rounding;
  * the own-leaf exclusion by index (softening below the FP32 range) against
    the oracle;
  * a group that falls back (list longer than the limit, forced through the
    environment) takes the union traversal of the "fp32" walk: bit-identical
    accelerations and counts;
  * on the 1M-body merger: at least as accurate as the per-body FP64 walk
    against the FP64 direct sum on a 4096-body subset, at theta 0.4, 0.5, 0.7;
  * the dispatch order changes no bit of a substep run (with fallbacks);
  * a device-resident substep run with the list walk tracks the FP64-strict
    trajectory, and the adjoint refuses the list precision.
"""
import os

import numpy as np
import pytest

import cosmosim_amd as ca
from oracle.cpu_ref import CpuRef
from _util import rel_err

pytestmark = pytest.mark.gpu
REF = CpuRef()

# FP32 list walk against the oracle's FP64 sums over the same lists: the
# offsets are differences of two FP32-rounded group-relative positions (error
# 2^-24 (|C - g| + |x - g|) per term), summed with compensation
TOL_RMS = 2e-6
TOL_MAX = 1e-4


def _sim(bodies, env=None, **kw):
    env = env or {}
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return ca.Simulation(bodies, **kw)
    finally:
        for k, v in saved.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


NO_FALLBACK = {"COSMOSIM_LIST_MAX": "1000000000"}


def _against_oracle(b, eps, theta):
    sim = _sim(b, NO_FALLBACK, softening=eps, theta=theta, precision="fp32_list")
    a = sim.accelerations()
    counts = sim.interaction_counts()
    st = sim.stats()
    _, wave, _ = sim.debug_orders()
    pool = sim.export_tree()
    sim.close()
    ra, rc, mass_err, bad = REF.list_walk(pool, b, wave, 1.0, eps, theta)
    return a, counts, st, ra, rc, mass_err, bad


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.8])
def test_list_walk_matches_oracle(theta):
    b = ca.generate(ca.IC_MERGER, 20000, 8)
    a, counts, st, ra, rc, mass_err, bad = _against_oracle(b, 0.01, theta)
    assert st.list_fallbacks == 0
    assert mass_err < 1e-12 and bad == 0
    assert np.array_equal(counts, rc)
    assert st.interactions == int(rc.sum())
    e = rel_err(a, ra)
    print(f"\ntheta {theta}: {rc.mean():.0f} entries per body, FP32 vs oracle rms {np.sqrt(np.mean(e**2)):.2e} max {e.max():.2e}")
    assert np.sqrt(np.mean(e ** 2)) < TOL_RMS and e.max() < TOL_MAX


def _cluster(seed, n=3000):
    """A collision run (300 bodies within 1e-9), exact duplicates (a depth-64 leaf) and
    dead bodies."""
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * np.array([2.0, 2.0, 0.2])
    pos[:300] = pos[0] + rng.normal(size=(300, 3)) * 1e-9
    pos[300:310] = pos[311]
    b = ca.make_bodies(pos, vel=np.zeros((n, 3)), mass=rng.uniform(0.1, 3.0, n))
    b["mass"][::41] = 0.0
    return b


def _uniform(n, seed):
    rng = np.random.default_rng(seed)
    return ca.make_bodies(rng.uniform(-1, 1, size=(n, 3)), vel=np.zeros((n, 3)),
                          mass=rng.uniform(0.5, 1.5, n))


EDGE = {
    "cluster": lambda: _cluster(1),
    "n1": lambda: _uniform(1, 3),
    "n2": lambda: _uniform(2, 4),
    "n65": lambda: _uniform(65, 5),
    "uniform_4096": lambda: _uniform(4096, 2),
}


# Each offset carries an absolute error of up to 2^-24 (|C - g| + |x - g|), set by
# the group's extent, not by the pair's distance.  In the cluster case 300 heavy
# bodies (masses to 3) lie within 1e-9 of each other, far below the softening
# length 0.01, and a group that holds some of them and some bodies 0.1 away puts
# ~6e-9 of error on offsets whose true size is 1e-9: each such pair's term
# G m dx / eps^3 is then off by G m 6e-9 / eps^3 ~ 1e-2 (the per-body walks form
# each pair's offset in FP64 and do not have this).
EDGE_TOL = {"cluster": (1e-3, 1e-2)}   # measured 3.3e-4 / 3.9e-3


@pytest.mark.parametrize("case", list(EDGE))
def test_list_walk_edge_inputs(case):
    b = EDGE[case]()
    a, counts, st, ra, rc, mass_err, bad = _against_oracle(b, 0.01, 0.5)
    live = ~(b["mass"] <= 0)
    assert st.list_fallbacks == 0 and mass_err < 1e-12 and bad == 0
    assert np.array_equal(counts, rc)
    assert np.all(a[~live] == 0.0)
    if live.sum() > 1:
        e = rel_err(a[live], ra[live])
        rms = np.sqrt(np.mean(e ** 2))
        print(f"\n{case}: FP32 vs oracle rms {rms:.2e} max {e.max():.2e}")
        tol_rms, tol_max = EDGE_TOL.get(case, (TOL_RMS, TOL_MAX))
        assert rms < tol_rms and e.max() < tol_max


@pytest.mark.parametrize("case,theta", [("merger", 0.3), ("merger", 0.5), ("cluster", 0.5),
                                        ("merger", 0.7)])
def test_total_from_list_lengths(case, theta):
    # the last walk of a substep run keeps only the interaction total; below
    # theta^2 = 1/3 it is taken from the list lengths (every own leaf is listed,
    # bodies merged into another's depth-64 leaf count it), above from per-entry
    # comparisons.  Either way it equals the per-body counts of a counting walk
    # on the same state, whose accelerations it matches bit for bit
    b = ca.generate(ca.IC_MERGER, 20000, 8) if case == "merger" else _cluster(2)
    sim = _sim(b, NO_FALLBACK, softening=0.01, theta=theta, dt=1e-4, precision="fp32_list")
    sim.advance(1)
    total, a1 = sim.stats().interactions, sim.accelerations()
    sim.compute_forces()
    counts = sim.interaction_counts()
    assert sim.stats().interactions == int(counts.sum()) == total
    assert np.array_equal(sim.accelerations(), a1)
    sim.close()


def test_list_walk_all_dead_and_empty():
    b = _uniform(100, 6)
    b["mass"][:] = 0.0
    sim = ca.Simulation(b, softening=0.01, theta=0.5, precision="fp32_list")
    sim.compute_forces()
    assert np.all(sim.accelerations() == 0.0)
    assert sim.stats().interactions == 0
    sim.advance(2)
    sim.close()


def test_list_walk_theta0_is_direct_sum():
    n = 3000
    b = ca.generate(ca.IC_DISK, n, 5)
    sim = _sim(b, NO_FALLBACK, softening=0.02, theta=0.0, precision="fp32_list")
    a = sim.accelerations()
    counts = sim.interaction_counts()
    assert sim.stats().list_fallbacks == 0
    d = sim.direct_forces(np.arange(n))
    sim.close()
    assert np.all(counts == n - 1)
    e = rel_err(a, d)
    assert np.sqrt(np.mean(e ** 2)) < TOL_RMS and e.max() < TOL_MAX


def test_list_walk_own_leaf_by_index():
    # softening^2 = 1e-14 is below the FP32 range the walk relies on for the own
    # leaf's zero term, so the own leaf is excluded by index.  Unsoftened close
    # pairs inside a group carry the group-relative rounding at its largest
    # (2^-24 times the group's extent over the pair's distance), so the bound is
    # looser than with softening
    b = ca.generate(ca.IC_MERGER, 8000, 9)
    a, counts, st, ra, rc, mass_err, bad = _against_oracle(b, 1e-7, 0.5)
    assert st.list_fallbacks == 0 and mass_err < 1e-12 and bad == 0
    assert np.array_equal(counts, rc)
    assert np.all(np.isfinite(a))
    e = rel_err(a, ra)
    print(f"\nunsoftened: FP32 vs oracle rms {np.sqrt(np.mean(e**2)):.2e} max {e.max():.2e}")
    assert np.sqrt(np.mean(e ** 2)) < 1e-5 and e.max() < 1e-3


ORDERED = {"COSMOSIM_HILBERT_MIN_WAVES": "0", "COSMOSIM_HILBERT_MIN_BODIES": "0"}


def test_fallback_is_the_union_walk():
    b = ca.generate(ca.IC_MERGER, 20000, 8)
    nw = (20000 + 63) // 64
    s1 = _sim(b, {"COSMOSIM_LIST_MAX": "0", "COSMOSIM_LIST_STICKY": "0", **ORDERED},
              softening=0.01, theta=0.5, precision="fp32_list")
    s1.set_timing(True)
    a1, c1 = s1.accelerations(), s1.interaction_counts()
    s1.compute_forces()
    assert np.array_equal(a1, s1.accelerations())
    assert s1.stats().list_fallbacks == nw
    s2 = _sim(b, ORDERED, softening=0.01, theta=0.5, precision="fp32")
    a2, c2 = s2.accelerations(), s2.interaction_counts()
    s1.close(); s2.close()
    assert np.array_equal(a1, a2)
    assert np.array_equal(c1, c2)


def test_fallen_back_waves_stay_on_the_union_walk():
    # a wave that fell back takes the union traversal in its next 3 walks (counted
    # as fallbacks only when the list is tried), then tries the list again; new
    # positions forget the history
    b = ca.generate(ca.IC_MERGER, 20000, 8)
    nw = (20000 + 63) // 64
    s = _sim(b, {"COSMOSIM_LIST_MAX": "0", "COSMOSIM_LIST_STICKY": "3", **ORDERED},
             softening=0.01, theta=0.5, precision="fp32_list")
    a0 = s.accelerations()          # the first walk (at create) fell back everywhere
    s.set_timing(True)
    seen = []
    for _ in range(5):
        s.compute_forces()
        seen.append(s.stats().list_fallbacks)
        assert np.array_equal(s.accelerations(), a0)
    assert seen == [0, 0, 0, nw, nw]
    s.set_state(pos=s.state()[0])
    s.set_timing(True)
    s.compute_forces()
    assert s.stats().list_fallbacks == nw
    s.close()


def test_dispatch_order_changes_no_result():
    # the dispatch order (longest first from each wave's and its neighbours' last
    # costs) only schedules the wavefronts: the same substeps without it give the
    # same bits, and so does a run whose fallbacks start from a small limit
    b = ca.generate(ca.IC_MERGER, 40000, 6)
    out = []
    for env in ({}, {"COSMOSIM_WAVE_ORDER": "0"}):
        s = _sim(b, {"COSMOSIM_LIST_MAX": "1500", **env}, softening=0.01, theta=0.5,
                 dt=1e-4, precision="fp32_list")
        s.set_timing(True)
        s.advance(4)
        out.append((s.accelerations(), s.state()[0], s.stats().list_fallbacks))
        s.close()
    assert out[0][2] > 0
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.fixture(scope="module")
def merger1m():
    b = ca.generate(ca.IC_MERGER, 1_000_000, 3)
    idx = np.sort(np.random.default_rng(0).choice(len(b), 4096, replace=False))
    return b, idx


@pytest.mark.parametrize("theta", [0.4, 0.5, 0.7])
def test_million_body_accuracy(merger1m, theta):
    b, idx = merger1m
    sim = ca.Simulation(b, softening=0.005, theta=theta, precision="fp32_list")
    a = sim.accelerations()
    nl = sim.stats().interactions
    direct = sim.direct_forces(idx)
    sim.set_params(precision="fp64")
    sim.compute_forces()
    a64 = sim.accelerations()
    n64 = sim.stats().interactions
    sim.close()
    e = rel_err(a[idx], direct)
    e64 = rel_err(a64[idx], direct)
    rms, rms64 = np.sqrt(np.mean(e ** 2)), np.sqrt(np.mean(e64 ** 2))
    print(f"\ntheta {theta}: list rms {rms:.3e} max {e.max():.3e}; per-body FP64 rms {rms64:.3e} "
          f"max {e64.max():.3e}; interactions {nl / 1e9:.2f}G vs {n64 / 1e9:.2f}G")
    assert nl > n64
    assert rms < rms64
    assert e.max() < 1.2 * e64.max()


def test_substeps_track_fp64_strict():
    b = ca.generate(ca.IC_MERGER, 100_000, 2)
    v0 = np.stack([b["vx"], b["vy"], b["vz"]], 1)
    s32 = ca.Simulation(b, softening=0.01, theta=0.5, dt=1e-3, precision="fp32_list")
    s64 = ca.Simulation(b, softening=0.01, theta=0.5, dt=1e-3, precision="fp64_strict")
    s32.advance(10)
    s64.advance(10)
    p32, v32 = s32.state()[:2]
    p64, v64 = s64.state()[:2]
    st = s32.stats()
    s32.close(); s64.close()
    assert st.n_live == 100_000 and st.interactions > 0
    # the velocity change over ten substeps is the time-averaged force, which the
    # two opening rules approximate to ~1e-2 each; positions move by |a| t^2 / 2
    dv = rel_err(v32 - v0, v64 - v0)
    print(f"\nvelocity change: rms rel diff {np.sqrt(np.mean(dv ** 2)):.2e}; "
          f"max |dx| {np.max(np.abs(p32 - p64)):.2e}")
    assert np.sqrt(np.mean(dv ** 2)) < 3e-2
    assert np.max(np.abs(p32 - p64)) < 1e-3


def test_adjoint_refuses_list_precision():
    b = ca.generate(ca.IC_MERGER, 2000, 4)
    sim = ca.Simulation(b, softening=0.02, theta=0.5, precision="fp32_list")
    with pytest.raises(ca.CosmosimError):
        sim.record(5)
    sim.close()
