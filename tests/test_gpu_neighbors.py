"""GPU: batched sphere queries over the device tree (cosmosim_ext_find_neighbors)
against octree_find_neighbors semantics (src/octree.c:236-318).

The lists must be identical, element for element: same bodies, same order
(children in slot order), same truncation at the cap.  Pinned by the reference's
own output (golden lists of the 4096-body disk) and by the oracle's restatement
on clustered inputs with a collision run, exact duplicates and dead bodies.
"""
import numpy as np
import pytest

import cosmosim_amd as ca
from oracle.cpu_ref import CpuRef
from tests._util import from_hot, golden

pytestmark = pytest.mark.gpu
ref = CpuRef()


def test_golden_lists_of_the_reference():
    g = golden("disk4096.npz")
    b = from_hot(g["bodies"])
    sim = ca.Simulation(b, softening=0.01, theta=0.5)
    offs = g["nb_offsets"]
    got = sim.find_neighbors(g["nb_centres"], float(g["nb_radius"]), int(g["nb_cap"]))
    for k in range(len(g["nb_centres"])):
        assert np.array_equal(got[k], g["nb_result"][offs[k]:offs[k + 1]])


def _cluster_bodies(seed, n=3000):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * np.array([2.0, 2.0, 0.2])
    pos[:300] = pos[0] + rng.normal(size=(300, 3)) * 1e-9     # one collision run
    pos[300:310] = pos[311]                                   # exact duplicates
    b = ca.make_bodies(pos, vel=rng.normal(size=(n, 3)) * 0.1, mass=rng.uniform(0.1, 3.0, n))
    b["mass"][::41] = 0.0
    return b


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_lists_with_collisions_dead_bodies_and_caps(seed):
    b = _cluster_bodies(seed)
    sim = ca.Simulation(b, softening=0.02, theta=0.5)
    pool = ref.build(b)
    rng = np.random.default_rng(seed)
    pos = np.stack([b["x"], b["y"], b["z"]], 1)
    q = rng.choice(len(b), 400, replace=False)
    centres = np.concatenate([pos[q], pos[q[:50]] + rng.normal(size=(50, 3)),   # bodies, off-body
                              [[40.0, 0.0, 0.0], [0.0, 0.0, 0.0]]])            # outside the root, origin
    radii = np.concatenate([rng.choice([1e-12, 0.05, 0.3, 1.0, 3.0], len(centres) - 2),
                            [0.5, 10.0]])
    for cap in [128, 16, 1, 0]:
        got = sim.find_neighbors(centres, radii, cap)
        for k, c in enumerate(centres):
            want = ref.find_neighbors(pool, b, c, float(radii[k]), cap)
            assert np.array_equal(got[k], want), (cap, k)
    # cap reached for the wide queries, and dead bodies never reported
    full = sim.find_neighbors(centres, radii, 16)
    assert any(len(f) == 16 for f in full)
    dead = set(np.nonzero(b["mass"] <= 0)[0])
    assert not any(dead & set(f.tolist()) for f in full)


def test_after_steps_queries_use_the_new_positions():
    b = ca.generate(ca.IC_DISK, 5000, 2)
    sim = ca.Simulation(b, softening=0.01, theta=0.5, dt=1e-3, precision="fp32")
    sim.advance(3)
    pos = sim.state()[0]
    b2 = b.copy()
    b2["x"], b2["y"], b2["z"] = pos[:, 0], pos[:, 1], pos[:, 2]
    pool = ref.build(b2)
    q = np.arange(0, 5000, 37)
    got = sim.find_neighbors(pos[q], 0.2, 64)
    for k, i in enumerate(q):
        assert np.array_equal(got[k], ref.find_neighbors(pool, b2, pos[i], 0.2, 64))


def test_million_body_queries_sampled_against_the_oracle():
    n = 1_000_000
    b = ca.generate(ca.IC_MERGER, n, 3)
    sim = ca.Simulation(b, softening=0.005, theta=0.5, precision="fp32")
    pos = np.stack([b["x"], b["y"], b["z"]], 1)
    q = np.random.default_rng(0).choice(n, 20000, replace=False)
    got = sim.find_neighbors(pos[q], 0.05, 128)
    counts = np.array([len(x) for x in got])
    assert counts.max() <= 128 and counts.min() >= 1          # every query finds itself
    for k in range(0, 20000, 97):                              # all within the radius, no repeats
        d = np.linalg.norm(pos[got[k]] - pos[q[k]], axis=1)
        assert np.all(d <= 0.05 * (1 + 1e-12)) and len(set(got[k].tolist())) == len(got[k])
    pool = sim.export_tree()
    for k in range(0, 20000, 499):
        assert np.array_equal(got[k], ref.find_neighbors(pool, b, pos[q[k]], 0.05, 128))
