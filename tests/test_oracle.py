"""Pins the CPU oracle (oracle/cpu_ref.c) - CPU only.

* against the golden fixtures written from the reference's compiled sources
  (tests/golden/gen_golden.py): bit-exact pools, accelerations, trajectories
  and neighbour lists;
* against the compiled reference directly on random inputs, where it was built
  (oracle/_ref; absent on machines without the reference tree);
* against what the mathematics fixes: the brute-force static definition of the
  tree (shallowest unshared cell path), closed-form masses / centres of mass, the
  theta = 0 walk equal to an independent numpy direct sum, and the reference
  tests' known answers (tests/test_physics.c:87-343, 396-421).
"""
import numpy as np
import pytest

from cosmosim_amd._abi import make_bodies
from oracle.cpu_ref import CpuRef, RefBuild
from tests._util import (acc, direct_numpy, energy_numpy, from_hot, golden, hot, rel_err,
                         same_bits, pools_same_structure, pools_identical)

ref = CpuRef()
KAT_NAMES = ["single", "two", "lattice", "two_x5", "newton3", "kat8", "dead3", "n1", "alldead",
             "coincident", "chain"]
FORCE_SETTINGS = [(0.0, 0.01), (0.5, 0.01), (0.0, 0.1), (0.0, 1e-4)]


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("name", KAT_NAMES)
def test_kat_pool_and_forces_match_reference(name):
    g = golden("kat.npz")
    b = from_hot(g[f"{name}_in"])
    pool = ref.build(b)
    gp = g[f"{name}_pool"]
    assert len(pool) == len(gp)
    for f in gp.dtype.names:
        assert np.array_equal(pool[f], gp[f]), f
    for th, eps in FORCE_SETTINGS:
        bb = b.copy()
        ref.forces(ref.build(bb), bb, 1.0, eps, th)
        assert same_bits(acc(bb), g[f"{name}_acc_t{th}_e{eps}"]), (th, eps)


def test_orbit_and_momentum_trajectories_match_reference():
    g = golden("kat.npz")
    b = from_hot(g["orbit_in"])
    ref.init_acc(b, 1.0, 0.01, 0.0)
    for _ in range(2000):
        ref.step(b, 1e-3, 1.0, 0.01, 0.0)
    assert same_bits(hot(b), g["orbit_out"])
    b = from_hot(g["mom4_in"])
    ref.init_acc(b, 1.0, 0.1, 0.0)
    for _ in range(500):
        ref.step(b, 2e-3, 1.0, 0.1, 0.0)
    assert same_bits(hot(b), g["mom4_out"])


def test_disk4096_matches_reference():
    g = golden("disk4096.npz")
    b = from_hot(g["bodies"])
    pool = ref.build(b)
    assert pools_identical(pool, g["pool"])
    for th in [0.0, 0.4, 0.5, 0.7]:
        bb = b.copy()
        ref.forces(pool, bb, 1.0, 0.01, th)
        assert same_bits(acc(bb), g[f"acc_t{th}"]), th


def test_neighbors_match_reference():
    g = golden("disk4096.npz")
    b = from_hot(g["bodies"])
    pool = ref.build(b)
    offs = g["nb_offsets"]
    for k, c in enumerate(g["nb_centres"]):
        got = ref.find_neighbors(pool, b, c, float(g["nb_radius"]), int(g["nb_cap"]))
        assert np.array_equal(got, g["nb_result"][offs[k]:offs[k + 1]])


def test_traj1024_matches_reference():
    g = golden("traj1024.npz")
    b = from_hot(g["bodies"])
    ref.init_acc(b, 1.0, 0.01, 0.5)
    assert same_bits(hot(b), g["init"])
    for _ in range(10):
        ref.step(b, 1e-3, 1.0, 0.01, 0.5)
    assert same_bits(hot(b), g["final"])


@pytest.mark.skipif(not RefBuild.available(), reason="reference build (oracle/_ref) not present")
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_equals_compiled_reference_random(seed):
    rng = np.random.default_rng(seed)
    n = 1500
    pos = rng.normal(size=(n, 3)) * np.array([2.0, 2.0, 0.2])
    pos[:40] = pos[0] + rng.normal(size=(40, 3)) * 1e-8      # a tight cluster
    pos[40:43] = pos[43]                                      # exact duplicates
    b = make_bodies(pos, vel=rng.normal(size=(n, 3)) * 0.1, mass=rng.uniform(0.1, 3.0, n))
    b["mass"][::53] = 0.0
    r = RefBuild()
    p1, p2 = ref.build(b), r.build(b)
    assert pools_identical(p1, p2)
    for th in [0.0, 0.3, 0.6]:
        b1, b2 = b.copy(), b.copy()
        ref.forces(p1, b1, 1.0, 0.02, th)
        r.forces(p2, b2, 1.0, 0.02, th)
        assert same_bits(hot(b1), hot(b2))
    b1, b2 = b.copy(), b.copy()
    ref.init_acc(b1, 1.0, 0.02, 0.5)
    r.init_acc(b2, 1.0, 0.02, 0.5)
    for _ in range(3):
        ref.step(b1, 1e-3, 1.0, 0.02, 0.5)
        r.step(b2, 1e-3, 1.0, 0.02, 0.5)
    assert same_bits(hot(b1), hot(b2))


# ------------------------------------------------------------ mathematics
def _paths(b, levels=64):
    """Octant path of every live body from the FP64 cell recurrence, computed
    here independently of the C code (src/octree.c:30-50, 141-159)."""
    live = [i for i in range(len(b)) if not (b["mass"][i] <= 0.0)]
    p = np.stack([b["x"], b["y"], b["z"]], 1)
    lo = p[live].min(axis=0)
    hi = p[live].max(axis=0)
    ext = max(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2])
    H = ext * 0.5 * 1.01
    if H < 1e-10:
        H = 1.0
    c0 = [(lo[a] + hi[a]) * 0.5 for a in range(3)]
    paths = {}
    for i in live:
        c, h, path = list(c0), H, []
        for _ in range(levels):
            q = (1 if p[i, 0] >= c[0] else 0) | (2 if p[i, 1] < c[1] else 0) | (4 if p[i, 2] < c[2] else 0)
            path.append(q)
            o = 0.5 * h
            c[0] = c[0] + o if q & 1 else c[0] - o
            c[1] = c[1] - o if q & 2 else c[1] + o
            c[2] = c[2] - o if q & 4 else c[2] + o
            h = h * 0.5
        paths[i] = tuple(path)
    return paths, H


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tree_equals_bruteforce_definition(seed):
    """A node exists for every path prefix shared by >= 2 live bodies (depth < 64)
    plus one leaf per body at its shallowest unshared prefix (depth 64 merges)."""
    rng = np.random.default_rng(seed)
    n = 60
    pos = rng.normal(size=(n, 3))
    pos[5:9] = pos[4] + rng.normal(size=(4, 3)) * 1e-6
    pos[10:12] = pos[12]
    b = make_bodies(pos, mass=rng.uniform(0.5, 2.0, n))
    b["mass"][::13] = 0.0
    paths, H = _paths(b)
    internal = set()
    for d in range(64):
        counts = {}
        for i, pth in paths.items():
            counts[pth[:d]] = counts.get(pth[:d], 0) + 1
        internal |= {k for k, v in counts.items() if v >= 2}
    leaf_depth = {}
    for i, pth in paths.items():
        d = 0
        while d < 64 and pth[:d] in internal:
            d += 1
        leaf_depth[i] = d
    merged = {}
    for i, pth in paths.items():
        merged.setdefault(pth[:leaf_depth[i]], []).append(i)
    pool = ref.build(b)
    assert len(pool) == len(internal) + len(merged)
    depth = np.rint(np.log2(H / pool["half_size"])).astype(int)
    leaves = pool[pool["body_index"] >= 0]
    ldep = depth[pool["body_index"] >= 0]
    assert sorted(leaves["body_index"]) == sorted(min(v) for v in merged.values())
    for node, d in zip(leaves, ldep):
        assert leaf_depth[int(node["body_index"])] == d
    int_depth = np.bincount(depth[pool["body_index"] < 0], minlength=65)
    want = np.bincount([len(k) for k in internal], minlength=65)
    assert np.array_equal(int_depth, want)


def test_masses_and_centres_closed_form():
    g = golden("disk4096.npz")
    b = from_hot(g["bodies"])
    pool = ref.build(b)
    live = b["mass"] > 0
    m = b["mass"][live]
    assert abs(pool[0]["total_mass"] - m.sum()) <= 1e-14 * m.sum()
    for a, f in enumerate(["x", "y", "z"]):
        com = np.sum(m * b[f][live]) / m.sum()
        assert abs(pool[0][["cx", "cy", "cz"][a]] - com) <= 1e-13
    leaves = pool[pool["body_index"] >= 0]
    bi = leaves["body_index"]
    assert np.array_equal(leaves["cx"], b["x"][bi]) and np.array_equal(leaves["total_mass"],
                                                                       b["mass"][bi])


def test_reference_known_answers():
    """tests/test_physics.c:87-343 and :396-421 restated on the oracle."""
    b = make_bodies([[3.0, 4.0, 5.0]], mass=[7.0])
    p = ref.build(b)
    assert abs(p[0]["total_mass"] - 7.0) < 1e-12 and abs(p[0]["cz"] - 5.0) < 1e-12
    b = make_bodies([[1, 0, 0], [-1, 0, 0]], mass=[1.0, 3.0])
    p = ref.build(b)
    assert abs(p[0]["total_mass"] - 4.0) < 1e-12 and abs(p[0]["cx"] + 0.5) < 1e-12
    b = make_bodies([[0, 0, 0], [5, 0, 0]], mass=[1.0, 1.0])
    ref.forces(ref.build(b), b, 1.0, 0.01, 0.0)
    assert b["ax"][0] > 0 and b["ax"][1] < 0 and abs(b["ay"][0]) < 1e-12
    b = make_bodies([[0, 0, 0], [3, 4, 0]], mass=[2.0, 5.0])
    ref.forces(ref.build(b), b, 1.0, 0.01, 0.0)
    assert abs(2.0 * b["ax"][0] + 5.0 * b["ax"][1]) < 1e-10
    r = []
    for d in (5.0, 10.0):
        b = make_bodies([[0, 0, 0], [d, 0, 0]], mass=[1.0, 1.0])
        ref.forces(ref.build(b), b, 1.0, 1e-4, 0.0)
        r.append(abs(b["ax"][0]))
    assert abs(r[0] / r[1] - 4.0) < 0.01
    b = make_bodies([[0, 0, 0], [5, 0, 0], [-5, 0, 0]], mass=[10.0, 0.0, 3.0])
    p = ref.build(b)
    assert abs(p[0]["total_mass"] - 13.0) < 1e-12
    ref.forces(p, b, 1.0, 0.01, 0.0)
    assert b["ax"][0] < 0 and b["ax"][1] == 0.0
    # 2-body orbit energy and 4-body momentum (golden inputs)
    g = golden("kat.npz")
    b = from_hot(g["orbit_in"])
    e0 = sum(energy_numpy(np.stack([b["x"], b["y"], b["z"]], 1),
                          np.stack([b["vx"], b["vy"], b["vz"]], 1), b["mass"], 1.0, 0.01))
    o = from_hot(g["orbit_out"])
    e1 = sum(energy_numpy(np.stack([o["x"], o["y"], o["z"]], 1),
                          np.stack([o["vx"], o["vy"], o["vz"]], 1), o["mass"], 1.0, 0.01))
    assert abs((e1 - e0) / e0) < 1e-2
    m0, m1 = from_hot(g["mom4_in"]), from_hot(g["mom4_out"])
    for f in ["vx", "vy", "vz"]:
        assert abs(np.sum(m0["mass"] * m0[f]) - np.sum(m1["mass"] * m1[f])) < 1e-10


def test_theta0_walk_equals_direct_sum():
    rng = np.random.default_rng(4)
    n = 400
    pos = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 2.0, n)
    b = make_bodies(pos, mass=m)
    ref.forces(ref.build(b), b, 1.0, 0.05, 0.0)
    want = direct_numpy(pos, m, np.arange(n), 1.0, 0.05)
    assert rel_err(acc(b), want).max() < 1e-12


def test_direct_and_energy_judges():
    rng = np.random.default_rng(5)
    n = 300
    pos, vel = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 2.0, n)
    b = make_bodies(pos, vel=vel, mass=m)
    idx = np.arange(0, n, 3)
    assert rel_err(ref.direct(b, idx, 1.0, 0.03), direct_numpy(pos, m, idx, 1.0, 0.03)).max() < 1e-12
    ke, pe = ref.energy(b, 1.0, 0.03)
    ke2, pe2 = energy_numpy(pos, vel, m, 1.0, 0.03)
    assert abs(ke - ke2) <= 1e-13 * abs(ke2) and abs(pe - pe2) <= 1e-12 * abs(pe2)


def test_canonical_renumbering_is_structure_preserving():
    g = golden("disk4096.npz")
    pool = g["pool"]
    c = ref.canonical(pool)
    assert len(c) == len(pool)
    assert pools_same_structure(ref.canonical(c), c)
    assert np.isclose(c[0]["total_mass"], pool[0]["total_mass"])
