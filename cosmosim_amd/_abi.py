"""Host-side mirrors of the C records in include/*.h.

``BODY_DTYPE`` / ``NODE_DTYPE`` are numpy structured dtypes with the exact byte
layout of ``Body`` (168 B, src/body.h:13-27) and ``OctreeNode`` (104 B,
src/octree.h:6-14); arrays of them are passed to the C-ABI as plain pointers.
The ctypes structures mirror ``CosmosimConfig`` (72 B, src/cosmosim_api.h:16-27)
and the extension records of include/cosmosim_ext.h.
"""
import ctypes

import numpy as np

_f8 = np.float64
_i4 = np.int32

BODY_FIELDS = [
    ("x", _f8, 0), ("y", _f8, 8), ("z", _f8, 16),
    ("vx", _f8, 24), ("vy", _f8, 32), ("vz", _f8, 40),
    ("ax", _f8, 48), ("ay", _f8, 56), ("az", _f8, 64),
    ("mass", _f8, 72), ("type", _i4, 80),
    ("spin_x", _f8, 88), ("spin_y", _f8, 96), ("spin_z", _f8, 104),
    ("accretion_rate", _f8, 112), ("luminosity", _f8, 120), ("lifetime", _f8, 128),
    ("density", _f8, 136), ("pressure", _f8, 144), ("internal_energy", _f8, 152),
    ("smoothing_h", _f8, 160),
]
BODY_DTYPE = np.dtype({
    "names": [f[0] for f in BODY_FIELDS],
    "formats": [f[1] for f in BODY_FIELDS],
    "offsets": [f[2] for f in BODY_FIELDS],
    "itemsize": 168,
})

NODE_DTYPE = np.dtype({
    "names": ["cx", "cy", "cz", "total_mass", "center_x", "center_y", "center_z",
              "half_size", "body_index", "children"],
    "formats": [_f8, _f8, _f8, _f8, _f8, _f8, _f8, _f8, _i4, (_i4, 8)],
    "offsets": [0, 8, 16, 24, 32, 40, 48, 56, 64, 68],
    "itemsize": 104,
})

BODY_STAR, BODY_GAS, BODY_SMBH, BODY_JET, BODY_DUST, BODY_LOBE = range(6)

PREC_FP64_STRICT = 0
PREC_FP64 = 1
PREC_FP32 = 2
PREC_FP32_LIST = 3
PRECISIONS = {"fp64_strict": PREC_FP64_STRICT, "fp64": PREC_FP64, "fp32": PREC_FP32,
              "fp32_list": PREC_FP32_LIST}

IC_DISK = 0
IC_MERGER = 1


class CosmosimConfig(ctypes.Structure):
    _fields_ = [
        ("n_bodies", ctypes.c_int),
        ("merger", ctypes.c_int),
        ("quasar", ctypes.c_int),
        ("dt", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("smbh_mass_frac", ctypes.c_double),
        ("accretion_radius", ctypes.c_double),
        ("jet_speed", ctypes.c_double),
        ("feedback_strength", ctypes.c_double),
        ("substeps", ctypes.c_int),
    ]


class ExtParams(ctypes.Structure):
    _fields_ = [
        ("G", ctypes.c_double),
        ("softening", ctypes.c_double),
        ("theta", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("device", ctypes.c_int),
    ]


class QuasarConfig(ctypes.Structure):
    """CosmosimQuasarConfig: the byte layout of the reference's QuasarConfig
    (src/quasar.h:6-22).  default() holds quasar_default_config's values
    (src/quasar.c:38-56)."""
    _fields_ = [(f, ctypes.c_double) for f in (
        "accretion_radius", "swallow_radius", "isco_radius", "viscosity_alpha", "jet_speed",
        "jet_mass", "jet_lifetime", "feedback_strength", "eta_eff", "eddington_k",
        "accretion_smoothing")] + [(f, ctypes.c_int) for f in (
            "jet_cap", "jet_count", "max_bodies", "jet_ring_count")]

    @classmethod
    def default(cls, **kw):
        c = cls(3.0, 0.3, 0.5, 0.05, 20.0, 0.5, 5.0, 1.0, 0.1, 0.5, 0.95, 2000, 0, 0, 4)
        for k, v in kw.items():
            setattr(c, k, v)
        return c


class ExtStats(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("n_live", ctypes.c_int),
        ("n_nodes", ctypes.c_int),
        ("max_depth", ctypes.c_int),
        ("collision_runs", ctypes.c_int),
        ("timed_walks", ctypes.c_int),
        ("interactions", ctypes.c_longlong),
        ("root_center", ctypes.c_double * 3),
        ("root_half", ctypes.c_double),
        ("ms_build", ctypes.c_double),
        ("ms_walk", ctypes.c_double),
        ("ms_kdk", ctypes.c_double),
        ("ms_neighbors", ctypes.c_double),
        ("neighbor_queries", ctypes.c_longlong),
        ("ms_sph", ctypes.c_double),
        ("list_fallbacks", ctypes.c_int),
        ("reserved", ctypes.c_int),
    ]


assert BODY_DTYPE.itemsize == 168 and NODE_DTYPE.itemsize == 104
assert ctypes.sizeof(CosmosimConfig) == 72


def make_bodies(pos, vel=None, mass=None, body_type=BODY_STAR):
    """Body array from (n,3) positions, optional (n,3) velocities and (n,) masses."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    b = np.zeros(n, dtype=BODY_DTYPE)
    b["x"], b["y"], b["z"] = pos[:, 0], pos[:, 1], pos[:, 2]
    if vel is not None:
        vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        b["vx"], b["vy"], b["vz"] = vel[:, 0], vel[:, 1], vel[:, 2]
    b["mass"] = 1.0 if mass is None else np.asarray(mass, dtype=np.float64)
    b["type"] = body_type
    return b


def positions(b):
    return np.stack([b["x"], b["y"], b["z"]], axis=1)


def velocities(b):
    return np.stack([b["vx"], b["vy"], b["vz"]], axis=1)


def accelerations(b):
    return np.stack([b["ax"], b["ay"], b["az"]], axis=1)


def ptr(a):
    """ctypes void pointer to a contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)
