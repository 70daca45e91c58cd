"""GPU: the FP32 walk's wavefronts take their bodies in Peano-Hilbert order.

The build derives each sorted body's Hilbert rank from the tree (per cell, its
children's Hilbert offsets minus their sorted offsets, summed over the enclosing
cells).  Checks, on inputs with collision runs, duplicates, dead bodies, a
depth-64 cell and tiny counts:

  * the wave order is a permutation of the live bodies;
  * it equals the stable sort of the sorted bodies by the Hilbert index of their
    48-bit keys, computed two ways on the host: the state table walked digit by
    digit, and Skilling's transform of the coordinates the key's bits spell
    (independent of the table);
  * the FP32 forces with Hilbert wavefronts equal those with sorted-order
    wavefronts to FP32 rounding, and the per-body interaction counts are equal:
    the grouping changes only which bodies share a wavefront;
  * the FP64 walks' forces do not depend on the grouping at all (bit-identical);
  * trees of fewer than 16384 bodies keep sorted-order wavefronts.

The library uses the order in the wave-mask walks only when there are more
wavefronts than the GPU holds at once; the tests lower that threshold (and the body
threshold) through the environment so that every input exercises the ordered walks.
"""
import os

import numpy as np
import pytest

import cosmosim_amd as ca
from test_hilbert_table import table

pytestmark = pytest.mark.gpu
T = table()


def hilbert_of_keys(keys):
    st = np.zeros(len(keys), np.int64)
    h = np.zeros(len(keys), np.int64)
    k = keys.astype(np.int64)
    for l in range(16):
        m = (k >> (3 * (15 - l))) & 7
        e = T[8 * st + m]
        h = (h << 3) | (e & 7)
        st = e >> 3
    return h


def skilling_of_keys(keys):
    """Skilling's transform, vectorised over keys: the 16-bit coordinates are the
    key's digit bits (bit 2 of each digit the first axis)."""
    k = keys.astype(np.int64)
    X = [np.zeros(len(k), np.int64) for _ in range(3)]
    for l in range(16):
        m = (k >> (3 * (15 - l))) & 7
        for a in range(3):
            X[a] = (X[a] << 1) | ((m >> (2 - a)) & 1)
    Q = 1 << 15
    while Q > 1:
        P = Q - 1
        for i in range(3):
            hit = (X[i] & Q) != 0
            X0 = np.where(hit, X[0] ^ P, X[0])
            t = np.where(hit, 0, (X[0] ^ X[i]) & P)
            X[0] = X0 ^ t
            if i:
                X[i] = X[i] ^ t
        Q >>= 1
    for i in range(1, 3):
        X[i] = X[i] ^ X[i - 1]
    t = np.zeros(len(k), np.int64)
    Q = 1 << 15
    while Q > 1:
        t = np.where((X[2] & Q) != 0, t ^ (Q - 1), t)
        Q >>= 1
    X = [x ^ t for x in X]
    h = np.zeros(len(k), np.int64)
    for b in range(15, -1, -1):
        for i in range(3):
            h = (h << 1) | ((X[i] >> b) & 1)
    return h


def cluster(seed, n=3000):
    rng = np.random.default_rng(seed)
    pos = rng.normal(size=(n, 3)) * np.array([2.0, 2.0, 0.2])
    pos[:300] = pos[0] + rng.normal(size=(300, 3)) * 1e-9     # one collision run
    pos[300:310] = pos[311]                                   # exact duplicates
    b = ca.make_bodies(pos, vel=np.zeros((n, 3)), mass=rng.uniform(0.1, 3.0, n))
    b["mass"][::41] = 0.0
    return b


def uniform(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, size=(n, 3))
    return ca.make_bodies(pos, vel=np.zeros((n, 3)), mass=rng.uniform(0.5, 1.5, n))


CASES = {
    "merger_100k": lambda: ca.generate(ca.IC_MERGER, 100_000, 3),
    "disk_20001": lambda: ca.generate(ca.IC_DISK, 20_001, 1),
    "cluster": lambda: cluster(1),
    "uniform_4096": lambda: uniform(4096, 2),
    "n1": lambda: uniform(1, 3),
    "n2": lambda: uniform(2, 4),
    "n65": lambda: uniform(65, 5),
}


def _sim(b, hilbert=True, min_bodies="1", min_waves="0"):
    """A handle whose builds compute the Peano-Hilbert order from min_bodies bodies and
    whose untaped walks use it (hilbert), the wave-mask walks from min_waves wavefronts
    (the defaults here: every input; the library's own defaults are 16384 bodies and
    the GPU's resident wavefronts)."""
    env = {"COSMOSIM_HILBERT_WAVES": "1" if hilbert else "0",
           "COSMOSIM_HILBERT_MIN_BODIES": min_bodies,
           "COSMOSIM_HILBERT_MIN_WAVES": min_waves}
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return ca.Simulation(b, softening=0.01, theta=0.5, dt=1e-3, precision="fp32")
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


@pytest.mark.parametrize("case", list(CASES))
def test_wave_order_is_the_hilbert_order_of_the_cells(case):
    b = CASES[case]()
    sim = _sim(b)
    sim.compute_forces()
    si, wi, keys = sim.debug_orders()
    live = np.nonzero(~(b["mass"] <= 0))[0]
    assert len(si) == len(live)
    assert np.array_equal(np.sort(wi), np.sort(si)) and np.array_equal(np.sort(si), live)
    h = hilbert_of_keys(keys)
    assert np.array_equal(h, skilling_of_keys(keys))
    perm = np.argsort(h, kind="stable")
    assert np.array_equal(wi, si[perm])
    sim.close()


@pytest.mark.parametrize("case", ["merger_100k", "cluster"])
def test_hilbert_wavefronts_change_only_the_grouping(case):
    b = CASES[case]()
    s1, s0 = _sim(b, True), _sim(b, False)
    s1.compute_forces()
    s0.compute_forces()
    a1, a0 = s1.accelerations(), s0.accelerations()
    assert np.array_equal(s1.interaction_counts(), s0.interaction_counts())
    live = ~(b["mass"] <= 0)
    scale = np.linalg.norm(a0[live], axis=1)
    err = np.linalg.norm(a1[live] - a0[live], axis=1) / np.maximum(scale, 1e-300)
    print(f"\n{case}: Hilbert vs sorted wavefronts, max relative difference {err.max():.2e}")
    assert err.max() <= 1e-5
    assert np.all(a1[~live] == 0.0)
    s1.close()
    s0.close()


@pytest.mark.parametrize("precision", ["fp64", "fp64_strict"])
def test_fp64_walks_do_not_depend_on_the_grouping(precision):
    """Each lane of an FP64 walk sums its own terms in FP64 (a lane that does not
    interact adds an exact zero), so Peano-Hilbert and sorted-order wavefronts give
    bit-identical accelerations and equal counts."""
    b = cluster(2)
    b2 = CASES["merger_100k"]()
    for bb in (b, b2):
        s1, s0 = _sim(bb, True), _sim(bb, False)
        for s in (s1, s0):
            s.set_params(precision=precision)
            s.compute_forces()
        assert np.array_equal(s1.accelerations().view(np.uint64), s0.accelerations().view(np.uint64))
        assert np.array_equal(s1.interaction_counts(), s0.interaction_counts())
        s1.close()
        s0.close()


def test_small_trees_keep_sorted_wavefronts():
    """With the library's thresholds: a build of fewer than 16384 bodies skips the
    order (the walks take sorted-order wavefronts); from 16384 bodies a build for the
    FP64 walks computes it."""
    for n, hil in ((4096, False), (16384, True)):
        b = uniform(n, 7)
        old = {k: os.environ.pop(k, None) for k in
               ("COSMOSIM_HILBERT_WAVES", "COSMOSIM_HILBERT_MIN_BODIES", "COSMOSIM_HILBERT_MIN_WAVES")}
        try:
            sim = ca.Simulation(b, softening=0.01, theta=0.5, dt=1e-3, precision="fp64")
        finally:
            os.environ.update({k: v for k, v in old.items() if v is not None})
        sim.compute_forces()
        si, wi, keys = sim.debug_orders()
        if hil:
            assert np.array_equal(wi, si[np.argsort(hilbert_of_keys(keys), kind="stable")])
            assert not np.array_equal(wi, si)
        else:
            assert np.array_equal(wi, si)
        sim.close()
