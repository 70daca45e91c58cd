"""Helpers shared by the tests: fixtures, body arrays, error measures."""
import os

import numpy as np

from cosmosim_amd._abi import BODY_DTYPE

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HOT = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "mass"]
STRUCT_FIELDS = ["center_x", "center_y", "center_z", "half_size", "body_index", "children"]


def golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def from_hot(h):
    """Body array from an [n, 10] array of x..az, mass."""
    b = np.zeros(len(h), dtype=BODY_DTYPE)
    for k, f in enumerate(HOT):
        b[f] = h[:, k]
    return b


def hot(b):
    return np.stack([b[f] for f in HOT], axis=1)


def acc(b):
    return np.stack([b["ax"], b["ay"], b["az"]], axis=1)


def rel_err(a, ref):
    """Per-body |a - ref| / |ref| (Euclidean; |ref| = 0 counts as 1)."""
    d = np.linalg.norm(np.asarray(a) - np.asarray(ref), axis=1)
    s = np.linalg.norm(ref, axis=1)
    s[s == 0] = 1.0
    return d / s


def same_bits(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def pools_identical(a, b):
    """Field-by-field equality of two pools (padding bytes ignored)."""
    return len(a) == len(b) and all(np.array_equal(a[f], b[f]) for f in a.dtype.names)


def pools_same_structure(canon_a, canon_b):
    """Canonical (pre-order) pools: identical cells, leaves and child slots."""
    if len(canon_a) != len(canon_b):
        return False
    return all(np.array_equal(canon_a[f], canon_b[f]) for f in STRUCT_FIELDS)


def pool_mass_com_err(canon_a, canon_b):
    m = np.max(np.abs(canon_a["total_mass"] - canon_b["total_mass"])
               / np.maximum(np.abs(canon_b["total_mass"]), 1e-300))
    scale = np.maximum(canon_b["half_size"], 1e-300)
    c = max(np.max(np.abs(canon_a[f] - canon_b[f]) / scale) for f in ["cx", "cy", "cz"])
    return m, c


def direct_numpy(pos, mass, idx, G, eps):
    """Independent vectorised FP64 direct sum (pairwise differences, numpy order)."""
    out = np.zeros((len(idx), 3))
    for t, i in enumerate(idx):
        d = pos - pos[i]
        r2 = np.einsum("ij,ij->i", d, d) + eps * eps
        w = G * mass / (r2 * np.sqrt(r2))
        w[i] = 0.0
        out[t] = (d * w[:, None]).sum(axis=0)
    return out


def energy_numpy(pos, vel, mass, G, eps):
    ke = 0.5 * np.sum(mass * np.einsum("ij,ij->i", vel, vel))
    pe = 0.0
    for i in range(len(pos) - 1):
        d = pos[i + 1:] - pos[i]
        r = np.sqrt(np.einsum("ij,ij->i", d, d) + eps * eps)
        pe -= G * mass[i] * np.sum(mass[i + 1:] / r)
    return ke, pe


BODY_FLOAT_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "mass", "spin_x", "spin_y",
                     "spin_z", "accretion_rate", "luminosity", "lifetime", "density", "pressure",
                     "internal_energy", "smoothing_h"]


def bodies_identical(a, b):
    """Every field of two body arrays bit-identical (padding ignored)."""
    if len(a) != len(b) or not np.array_equal(a["type"], b["type"]):
        return False
    return all(same_bits(a[f], b[f]) for f in BODY_FLOAT_FIELDS)


def first_body_difference(a, b):
    for f in BODY_FLOAT_FIELDS:
        d = np.nonzero(np.asarray(a[f]).view(np.uint64) != np.asarray(b[f]).view(np.uint64))[0]
        if len(d):
            return f, int(d[0]), float(a[f][d[0]]), float(b[f][d[0]])
    return None
