"""CPU: the oracle's restatement of the interaction-list walk (cref_list_walk).

The list walk has no counterpart in the reference; these checks pin the
restatement by what the mathematics fixes:

  * theta = 0 lists every body, so the sums are the FP64 direct sum
    (cref_direct) and every body sums n - 1 entries;
  * every group's list covers the live mass exactly once, and every listed cell
    passes the reference criterion (2h)^2/d^2 < theta^2 for every body of its
    group (the restatement counts both);
  * each body's list refines its per-body interaction set (a listed cell passes
    the body's own test, so the per-body walk takes that cell or an ancestor):
    at least as many entries as the reference walk's interactions for every
    body, and a smaller RMS error against the direct sum;
  * a group of one body is the per-body walk: the same interactions and counts.
"""
import numpy as np
import pytest

import cosmosim_amd as ca
from oracle.cpu_ref import CpuRef
from _util import acc, rel_err

REF = CpuRef()


def leaf_order(pool):
    """Bodies in the tree's depth-first order (children in slot order)."""
    out, stack = [], [0]
    while stack:
        nd = pool[stack.pop()]
        if nd["body_index"] >= 0:
            out.append(int(nd["body_index"]))
            continue
        for c in nd["children"][::-1]:
            if c >= 0:
                stack.append(int(c))
    return np.array(out, dtype=np.int32)


@pytest.fixture(scope="module")
def merger():
    b = ca.generate(ca.IC_MERGER, 3000, 11)
    pool = REF.build(b)
    return b, pool, leaf_order(pool)


def test_theta0_is_the_direct_sum(merger):
    b, pool, order = merger
    a, counts, mass_err, bad = REF.list_walk(pool, b, order, 1.0, 0.01, 0.0)
    d = REF.direct(b, np.arange(len(b)), 1.0, 0.01)
    assert np.all(counts == len(b) - 1)
    assert rel_err(a, d).max() < 1e-12
    assert mass_err < 1e-13 and bad == 0


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7, 1.0])
@pytest.mark.parametrize("shuffled", [False, True])
def test_lists_cover_the_mass_and_pass_the_criterion(merger, theta, shuffled):
    b, pool, order = merger
    if shuffled:   # wide groups: long lists, the same invariants
        order = np.random.default_rng(1).permutation(order).astype(np.int32)
    _, counts, mass_err, bad = REF.list_walk(pool, b, order, 1.0, 0.01, theta)
    assert mass_err < 1e-12
    assert bad == 0
    assert counts.min() > 0


@pytest.mark.parametrize("theta", [0.4, 0.5, 0.7])
def test_lists_refine_the_per_body_walk(merger, theta):
    b, pool, order = merger
    a, counts, _, _ = REF.list_walk(pool, b, order, 1.0, 0.01, theta)
    bb = b.copy()
    per_body = REF.forces(pool, bb, 1.0, 0.01, theta)
    assert np.all(counts >= per_body)
    d = REF.direct(b, np.arange(len(b)), 1.0, 0.01)
    e_list, e_body = rel_err(a, d), rel_err(acc(bb), d)
    assert np.sqrt(np.mean(e_list ** 2)) < np.sqrt(np.mean(e_body ** 2))


def test_groups_of_one_body_are_the_per_body_walk():
    # a group of one body has a point box: the group test is the body's own test
    # (up to the FP32 evaluation of the distance); with bodies 64 apart in the
    # order, every group holds one body when the order has one body per group
    b = ca.generate(ca.IC_DISK, 64, 3)
    pool = REF.build(b)
    bb = b.copy()
    per_body = REF.forces(pool, bb, 1.0, 0.05, 0.6)
    for i in range(len(b)):
        a, counts, mass_err, bad = REF.list_walk(pool, b, np.array([i], np.int32), 1.0, 0.05, 0.6)
        assert counts[i] == per_body[i]
        assert rel_err(a[i:i + 1], acc(bb)[i:i + 1])[0] < 1e-12
        assert mass_err < 1e-13 and bad == 0
