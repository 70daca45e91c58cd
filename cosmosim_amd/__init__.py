"""cosmosim_amd - MI355X-native gravity core for cosmosim (octree build,
Barnes-Hut walk, KDK leapfrog) behind the reference's C entry points.

The product is the shared library ``cosmosim_amd/libcosmosim.so`` (HIP kernels
for gfx950 + a C host layer).  This module is a thin ctypes mirror of its C-ABI
with the reference's names and argument meanings, so tests read like the
reference's own (tests/test_physics.c).  There is no Python or CPU fallback:
every call goes through the library, and a missing library or device raises.
"""
import ctypes
import os

import numpy as np

from ._abi import (BODY_DTYPE, NODE_DTYPE, CosmosimConfig, ExtParams, ExtStats, QuasarConfig,
                   PRECISIONS, PREC_FP64_STRICT, PREC_FP64, PREC_FP32, PREC_FP32_LIST, IC_DISK, IC_MERGER,
                   make_bodies, positions, velocities, accelerations, ptr)

__all__ = [
    "BODY_DTYPE", "NODE_DTYPE", "CosmosimConfig", "ExtParams", "ExtStats", "QuasarConfig",
    "PRECISIONS", "PREC_FP64_STRICT", "PREC_FP64", "PREC_FP32", "PREC_FP32_LIST", "IC_DISK", "IC_MERGER", "make_bodies",
    "positions", "velocities", "accelerations", "lib", "library_path", "CosmosimError",
    "octree_build", "octree_compute_forces", "octree_find_neighbors", "integrator_step",
    "integrator_init_accelerations", "generate", "default_orbit", "orbit_coefficients",
    "Simulation", "Handle", "device_info", "device_memory", "EXPORTED_SYMBOLS",
]

library_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcosmosim.so")

_VP = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int

# name -> (restype, argtypes) of every function declared in include/*.h
_SIGS = {
    # cosmosim_api.h
    "cosmosim_default_config": (CosmosimConfig, []),
    "cosmosim_create": (_VP, [CosmosimConfig]),
    "cosmosim_step": (None, [_VP]),
    "cosmosim_get_bodies": (_VP, [_VP]),
    "cosmosim_get_count": (_I, [_VP]),
    "cosmosim_get_active_count": (_I, [_VP]),
    "cosmosim_get_sim_time": (_D, [_VP]),
    "cosmosim_destroy": (None, [_VP]),
    # octree.h / integrator.h
    "octree_build": (None, [_VP, _VP, _VP, _I]),
    "octree_compute_forces": (None, [_VP, _I, _VP, _I, _D, _D, _D]),
    "octree_find_neighbors": (_I, [_VP, _I, _VP, _I, _D, _D, _D, _D, _VP, _I]),
    "integrator_step": (None, [_VP, _I, _D, _D, _D, _D, _VP, _I]),
    "integrator_init_accelerations": (None, [_VP, _I, _D, _D, _D, _VP]),
    # cosmosim_ext.h
    "cosmosim_ext_generate": (_I, [_I, _I, ctypes.c_ulonglong, _VP, _VP]),
    "cosmosim_ext_default_orbit": (_I, [_VP]),
    "cosmosim_ext_orbit_coefficients": (_I, [_I, _I, _VP]),
    "cosmosim_ext_create": (_I, [_VP, _I, ctypes.POINTER(ExtParams), ctypes.POINTER(_VP)]),
    "cosmosim_ext_set_params": (_I, [_VP, ctypes.POINTER(ExtParams)]),
    "cosmosim_ext_get_params": (_I, [_VP, ctypes.POINTER(ExtParams)]),
    "cosmosim_ext_step": (_I, [_VP, _I]),
    "cosmosim_ext_sync": (_I, [_VP]),
    "cosmosim_ext_upload": (_I, [_VP]),
    "cosmosim_ext_compute_forces": (_I, [_VP]),
    "cosmosim_ext_get_state": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "cosmosim_ext_set_state": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "cosmosim_ext_direct_forces": (_I, [_VP, _VP, _I, _VP]),
    "cosmosim_ext_energy": (_I, [_VP, _VP, _VP, _VP]),
    "cosmosim_ext_find_neighbors": (_I, [_VP, _VP, _VP, _I, _I, _VP, _VP]),
    "cosmosim_ext_sph_density": (_I, [_VP]),
    "cosmosim_ext_sph_forces": (_I, [_VP]),
    "cosmosim_ext_sph_cooling": (_I, [_VP, _D]),
    "cosmosim_ext_step_sph": (_I, [_VP, _I]),
    "cosmosim_ext_snapshot_ue": (_I, [_VP, _VP, _I, _VP]),
    "cosmosim_ext_snapshot_render": (_I, [_VP, _VP, _VP]),
    "cosmosim_ext_reserve": (_I, [_VP, _I]),
    "cosmosim_ext_quasar_step": (_I, [_VP, _VP, _D]),
    "cosmosim_ext_quasar_compact": (_I, [_VP, _VP]),
    "cosmosim_ext_quasar_state": (_I, [_VP, _VP, _VP, _I]),
    "cosmosim_ext_quasar_config": (_I, [_VP, _VP, _I]),
    "cosmosim_ext_interaction_counts": (_I, [_VP, _VP, _I]),
    "cosmosim_ext_get_stats": (_I, [_VP, ctypes.POINTER(ExtStats)]),
    "cosmosim_ext_set_timing": (_I, [_VP, _I]),
    "cosmosim_ext_export_tree": (_I, [_VP, _VP, _I, _VP]),
    "cosmosim_ext_handle_error": (_I, [_VP]),
    "cosmosim_ext_debug_fail_build": (_I, [_VP, _I]),
    "cosmosim_ext_debug_orders": (_I, [_VP, _VP, _VP, _VP, _I, _VP]),
    "cosmosim_ext_last_error": (_I, []),
    "cosmosim_ext_clear_error": (None, []),
    "cosmosim_ext_error_string": (ctypes.c_char_p, [_I]),
    "cosmosim_ext_forces_noalloc": (_I, [_VP, _VP, _VP, _VP, _I, _D, _D, _D, _I, _VP, _VP,
                                         _VP, _VP, _VP]),
    "cosmosim_ext_device_info": (_I, [_I, ctypes.c_char_p, _I, _VP]),
    "cosmosim_ext_device_memory": (_I, [_I, _VP, _VP]),
    "cosmosim_ext_force_vjp": (_I, [_VP, _VP, _VP]),
    "cosmosim_ext_forces_frozen": (_I, [_VP, _VP, _VP]),
    "cosmosim_ext_direct_vjp": (_I, [_VP, _VP, _VP]),
    "cosmosim_ext_record": (_I, [_VP, _I, _I, _I]),
    "cosmosim_ext_backward": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP]),
    "cosmosim_ext_replay": (_I, [_VP, _VP, _VP, _VP, _VP]),
    # csg.h (device-core C-ABI)
    "csg_device_count": (_I, [_VP]),
    "csg_device_info": (_I, [_I, ctypes.c_char_p, _I, _VP]),
    "csg_device_memory": (_I, [_I, _VP, _VP]),
    "csg_create": (_I, [_I, _I, _VP]),
    "csg_destroy": (None, [_VP]),
    "csg_reserve": (_I, [_VP, _I]),
    "csg_size": (_I, [_VP]),
    "csg_set_bodies": (_I, [_VP, _VP, _I]),
    "csg_get_bodies": (_I, [_VP, _VP, _I]),
    "csg_set_accelerations": (_I, [_VP, _VP, _I]),
    "csg_get_state": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "csg_set_state": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "csg_kick_drift": (_I, [_VP, _D]),
    "csg_kick": (_I, [_VP, _D]),
    "csg_build": (_I, [_VP]),
    "csg_walk": (_I, [_VP, _I, _D, _D, _D]),
    "csg_substeps": (_I, [_VP, _I, _D, _D, _D, _D, _I]),
    "csg_export_pool": (_I, [_VP, _VP, _I, _VP]),
    "csg_walk_pool": (_I, [_VP, _VP, _I, _I, _D, _D, _D]),
    "csg_direct_forces": (_I, [_VP, _VP, _I, _D, _D, _VP]),
    "csg_energy": (_I, [_VP, _D, _D, _VP, _VP, _VP]),
    "csg_find_neighbors": (_I, [_VP, _VP, _VP, _I, _I, _VP, _VP]),
    "csg_sph_density": (_I, [_VP]),
    "csg_sph_forces": (_I, [_VP]),
    "csg_sph_cooling": (_I, [_VP, _D]),
    "csg_substeps_sph": (_I, [_VP, _I, _D, _D, _D, _D, _I]),
    "csg_snapshot": (_I, [_VP, _I, _VP, _I, _VP]),
    "csg_quasar_step": (_I, [_VP, _VP, _D]),
    "csg_quasar_compact": (_I, [_VP, _VP]),
    "csg_quasar_state": (_I, [_VP, _VP, _VP, _I]),
    "csg_interaction_counts": (_I, [_VP, _VP, _I]),
    "csg_get_stats": (_I, [_VP, _VP]),
    "csg_set_timing": (_I, [_VP, _I]),
    "csg_synchronize": (_I, [_VP]),
    "csg_debug_fail_build": (_I, [_VP, _I]),
    "csg_debug_orders": (_I, [_VP, _VP, _VP, _VP, _I, _VP]),
    "csg_walk_taped": (_I, [_VP, _I, _D, _D, _D]),
    "csg_force_vjp": (_I, [_VP, _VP, _VP]),
    "csg_forces_frozen": (_I, [_VP, _VP, _VP]),
    "csg_direct_vjp": (_I, [_VP, _VP, _VP, _D, _D]),
    "csg_record": (_I, [_VP, _I, _D, _D, _D, _D, _I, _I, _I]),
    "csg_backward": (_I, [_VP, _VP, _VP, _I, _VP, _VP, _VP]),
    "csg_replay": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "csg_set_tape": (_I, [_VP, _I]),
    "csg_host_alloc": (_VP, [ctypes.c_size_t]),
    "csg_host_free": (None, [_VP]),
}
EXPORTED_SYMBOLS = tuple(_SIGS)

_libs = {}


class CosmosimError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        msg = lib().cosmosim_ext_error_string(code).decode() if _libs else str(code)
        super().__init__(f"{where}: {msg} (code {code})")


def lib(path=None):
    """The loaded libcosmosim.so, or another build of it at `path` (a variant
    linked with the caller's initial-condition generators); raises if missing."""
    path = library_path if path is None else os.path.abspath(path)
    if path not in _libs:
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it with `python -m cosmosim_amd.build`")
        L = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _libs[path] = L
    return _libs[path]


def _check(code, where):
    if code != 0:
        raise CosmosimError(code, where)


def _legacy_check(where):
    code = lib().cosmosim_ext_last_error()
    if code != 0:
        lib().cosmosim_ext_clear_error()
        raise CosmosimError(code, where)


def device_info(device=0):
    name = ctypes.create_string_buffer(256)
    cu = ctypes.c_int(0)
    _check(lib().cosmosim_ext_device_info(device, name, 256, ctypes.byref(cu)), "device_info")
    return name.value.decode(), cu.value


def device_memory(device=0):
    """(free, total) device memory in bytes."""
    f, t = ctypes.c_ulonglong(0), ctypes.c_ulonglong(0)
    _check(lib().cosmosim_ext_device_memory(device, ctypes.byref(f), ctypes.byref(t)),
           "device_memory")
    return f.value, t.value


# ------------------------------------------------------------ legacy (src/octree.h)
def octree_build(bodies):
    """Tree of the live bodies in reference layout (root at 0); returns the used pool."""
    n = len(bodies)
    pool = np.zeros(max(8 * max(n, 1), 16), dtype=NODE_DTYPE)
    used = ctypes.c_int(0)
    lib().octree_build(ptr(pool), ctypes.byref(used), ptr(bodies), n)
    _legacy_check("octree_build")
    return pool[: used.value].copy(), pool


def octree_compute_forces(pool, bodies, G, softening, theta, root=0):
    lib().octree_compute_forces(ptr(pool), root, ptr(bodies), len(bodies), G, softening, theta)
    _legacy_check("octree_compute_forces")


def octree_find_neighbors(pool, bodies, center, radius, capacity, root=0):
    buf = np.zeros(max(capacity, 1), dtype=np.int32)
    k = lib().octree_find_neighbors(ptr(pool), root, ptr(bodies), len(bodies), center[0],
                                    center[1], center[2], radius, ptr(buf), capacity)
    return buf[:k].copy()


def integrator_init_accelerations(bodies, G, softening, theta, pool=None):
    if pool is None:
        pool = np.zeros(max(8 * len(bodies), 16), dtype=NODE_DTYPE)
    lib().integrator_init_accelerations(ptr(bodies), len(bodies), G, softening, theta, ptr(pool))
    _legacy_check("integrator_init_accelerations")
    return pool


def integrator_step(bodies, dt, G, softening, theta, pool=None, sph_enabled=0):
    if pool is None:
        pool = np.zeros(max(8 * len(bodies), 16), dtype=NODE_DTYPE)
    lib().integrator_step(ptr(bodies), len(bodies), dt, G, softening, theta, ptr(pool),
                          sph_enabled)
    _legacy_check("integrator_step")
    return pool


# ------------------------------------------------------------ initial conditions
def generate(recipe, n, seed, orbit=None):
    """Seeded bodies: recipe IC_DISK or IC_MERGER, n bodies including SMBHs."""
    out = np.zeros(n, dtype=BODY_DTYPE)
    o = None if orbit is None else np.ascontiguousarray(orbit, dtype=np.float64)
    _check(lib().cosmosim_ext_generate(recipe, n, seed, ptr(o), ptr(out)), "generate")
    return out


def default_orbit():
    o = np.zeros(6)
    _check(lib().cosmosim_ext_default_orbit(ptr(o)), "default_orbit")
    return o


def orbit_coefficients(n, recipe=IC_MERGER):
    c = np.zeros(n)
    _check(lib().cosmosim_ext_orbit_coefficients(recipe, n, ptr(c)), "orbit_coefficients")
    return c


# ------------------------------------------------------------ handles
class Handle:
    """The eight legacy exports of src/cosmosim_api.h on one handle."""

    def __init__(self, config=None, library=None, **kw):
        self.L = L = lib(library)
        cfg = L.cosmosim_default_config() if config is None else config
        for k, v in kw.items():
            setattr(cfg, k, v)
        self.config = cfg
        self.h = L.cosmosim_create(cfg)
        if not self.h:
            raise CosmosimError(L.cosmosim_ext_last_error() or -3, "cosmosim_create")

    def quasar_state(self, rng=None, phase=None):
        """Get [rng, phase] of the quasar random stream; or set them when given."""
        r, p = ctypes.c_uint64(0), ctypes.c_int(0)
        if rng is None and phase is None:
            _check(self.L.cosmosim_ext_quasar_state(self.h, ctypes.byref(r), ctypes.byref(p), 0),
                   "quasar_state")
            return [r.value, p.value]
        r.value, p.value = int(rng), int(phase)
        _check(self.L.cosmosim_ext_quasar_state(self.h, ctypes.byref(r), ctypes.byref(p), 1),
               "quasar_state")
        return [r.value, p.value]

    def quasar_config(self, cfg=None):
        """The QuasarConfig the quasar-mode step uses (a copy); or set it."""
        if cfg is None:
            cfg = QuasarConfig()
            _check(self.L.cosmosim_ext_quasar_config(self.h, ctypes.byref(cfg), 0),
                   "quasar_config")
            return cfg
        _check(self.L.cosmosim_ext_quasar_config(self.h, ctypes.byref(cfg), 1), "quasar_config")
        return cfg

    def step(self):
        self.L.cosmosim_step(self.h)
        _check(self.L.cosmosim_ext_handle_error(self.h), "cosmosim_step")

    def bodies(self):
        n = self.count()
        p = self.L.cosmosim_get_bodies(self.h)
        buf = (ctypes.c_char * (168 * n)).from_address(p)
        return np.frombuffer(buf, dtype=BODY_DTYPE, count=n).copy()

    def count(self):
        return self.L.cosmosim_get_count(self.h)

    def active_count(self):
        return self.L.cosmosim_get_active_count(self.h)

    def sim_time(self):
        return self.L.cosmosim_get_sim_time(self.h)

    def close(self):
        if self.h:
            self.L.cosmosim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation(Handle):
    """Extension handle: device-resident stepping of caller-supplied bodies."""

    def __init__(self, bodies, G=1.0, softening=0.01, theta=0.5, dt=1e-3, precision="fp64",
                 device=0):
        self.L = L = lib()
        self.params = ExtParams(G, softening, theta, dt, PRECISIONS.get(precision, precision),
                                device)
        self.n = len(bodies)
        h = ctypes.c_void_p()
        b = np.ascontiguousarray(bodies, dtype=BODY_DTYPE)
        _check(L.cosmosim_ext_create(ptr(b), self.n, ctypes.byref(self.params), ctypes.byref(h)),
               "cosmosim_ext_create")
        self.h = h.value

    def set_params(self, **kw):
        for k, v in kw.items():
            if k == "precision":
                v = PRECISIONS.get(v, v)
            setattr(self.params, k, v)
        _check(self.L.cosmosim_ext_set_params(self.h, ctypes.byref(self.params)), "set_params")

    def advance(self, substeps):
        _check(self.L.cosmosim_ext_step(self.h, substeps), "cosmosim_ext_step")

    def compute_forces(self):
        _check(self.L.cosmosim_ext_compute_forces(self.h), "compute_forces")

    def state(self):
        pos, vel, acc = (np.zeros((self.n, 3)) for _ in range(3))
        mass = np.zeros(self.n)
        _check(self.L.cosmosim_ext_get_state(self.h, ptr(pos), ptr(vel), ptr(acc), ptr(mass)),
               "get_state")
        return pos, vel, acc, mass

    def set_state(self, pos=None, vel=None, acc=None, mass=None):
        arr = [None if a is None else np.ascontiguousarray(a, dtype=np.float64)
               for a in (pos, vel, acc, mass)]
        _check(self.L.cosmosim_ext_set_state(self.h, *(ptr(a) for a in arr)), "set_state")

    def accelerations(self):
        acc = np.zeros((self.n, 3))
        _check(self.L.cosmosim_ext_get_state(self.h, None, None, ptr(acc), None), "get_state")
        return acc

    def direct_forces(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.zeros((len(idx), 3))
        _check(self.L.cosmosim_ext_direct_forces(self.h, ptr(idx), len(idx), ptr(acc)),
               "direct_forces")
        return acc

    def energy(self):
        ke, pe = ctypes.c_double(), ctypes.c_double()
        p = np.zeros(3)
        _check(self.L.cosmosim_ext_energy(self.h, ctypes.byref(ke), ctypes.byref(pe), ptr(p)),
               "energy")
        return ke.value, pe.value, p

    def interaction_counts(self):
        c = np.zeros(self.n, dtype=np.int32)
        _check(self.L.cosmosim_ext_interaction_counts(self.h, ptr(c), self.n), "counts")
        return c

    def stats(self):
        s = ExtStats()
        _check(self.L.cosmosim_ext_get_stats(self.h, ctypes.byref(s)), "stats")
        return s

    def set_timing(self, on=True):
        _check(self.L.cosmosim_ext_set_timing(self.h, int(on)), "set_timing")

    def find_neighbors(self, centres, radius, cap=128):
        """Sphere queries over the device tree: a list of index arrays, one per
        centre, equal to octree_find_neighbors' output for each (radius: scalar or
        one per centre)."""
        centres = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 3)
        k = len(centres)
        rad = np.ascontiguousarray(np.broadcast_to(np.asarray(radius, dtype=np.float64), (k,)))
        out = np.zeros((k, max(cap, 1)), dtype=np.int32)
        counts = np.zeros(k, dtype=np.int32)
        _check(self.L.cosmosim_ext_find_neighbors(self.h, ptr(centres), ptr(rad), k, cap, ptr(out),
                                                 ptr(counts)), "find_neighbors")
        return [out[q, :counts[q]].copy() for q in range(k)]

    # ------------------------------------------------------------ SPH
    def sph_density(self):
        """sph_compute_density on the device (src/sph.c:45-95)."""
        _check(self.L.cosmosim_ext_sph_density(self.h), "sph_density")

    def sph_forces(self):
        """sph_compute_forces on the device (src/sph.c:99-196): adds to a."""
        _check(self.L.cosmosim_ext_sph_forces(self.h), "sph_forces")

    def sph_cooling(self, dt):
        """sph_apply_cooling on the device (src/sph.c:200-231)."""
        _check(self.L.cosmosim_ext_sph_cooling(self.h, dt), "sph_cooling")

    def advance_sph(self, substeps):
        """Substeps of integrator_step with SPH enabled, device-resident."""
        _check(self.L.cosmosim_ext_step_sph(self.h, substeps), "cosmosim_ext_step_sph")

    # ------------------------------------------------------------ quasar
    def reserve(self, capacity):
        """Grow the body capacity (room for bodies a quasar step appends)."""
        _check(self.L.cosmosim_ext_reserve(self.h, int(capacity)), "reserve")

    def quasar_step(self, cfg, dt):
        """quasar_step on the device (src/quasar.c:477-510); cfg (QuasarConfig) is
        updated in place; returns the new body count."""
        _check(self.L.cosmosim_ext_quasar_step(self.h, ctypes.byref(cfg), dt), "quasar_step")
        self.n = self.count()
        return self.n

    def quasar_compact(self):
        """quasar_compact on the device (src/quasar.c:512-523); returns the new count."""
        k = ctypes.c_int(0)
        _check(self.L.cosmosim_ext_quasar_compact(self.h, ctypes.byref(k)), "quasar_compact")
        self.n = self.count()
        return k.value

    # ------------------------------------------------------------ snapshots
    def snapshot_ue(self, max_bodies=None):
        """FCosmosimBodyGPU records of the live bodies: float32 [count, 16]."""
        m = self.n if max_bodies is None else int(max_bodies)
        out = np.zeros((max(m, 1), 16), dtype=np.float32)
        k = ctypes.c_int(0)
        _check(self.L.cosmosim_ext_snapshot_ue(self.h, ptr(out), m, ctypes.byref(k)), "snapshot_ue")
        return out[: k.value].copy()

    def snapshot_render(self):
        """The renderer's upload records of the live bodies: float32 [count, 10]."""
        out = np.zeros((max(self.n, 1), 10), dtype=np.float32)
        k = ctypes.c_int(0)
        _check(self.L.cosmosim_ext_snapshot_render(self.h, ptr(out), ctypes.byref(k)),
               "snapshot_render")
        return out[: k.value].copy()

    def export_tree(self):
        cap = 8 * self.n + 4096
        pool = np.zeros(cap, dtype=NODE_DTYPE)
        used = ctypes.c_int(0)
        _check(self.L.cosmosim_ext_export_tree(self.h, ptr(pool), cap, ctypes.byref(used)),
               "export_tree")
        return pool[: used.value].copy()

    def sync(self):
        _check(self.L.cosmosim_ext_sync(self.h), "sync")

    def debug_fail_build(self, nth):
        """Test hook: the nth tree build from now fails as an over-capacity tree."""
        _check(self.L.cosmosim_ext_debug_fail_build(self.h, int(nth)), "debug_fail_build")

    def debug_orders(self):
        """Test hook: (sorted_idx, wave_idx, keys) of the last tree, n_live entries each:
        the bodies in key order, the FP32 walk's wavefront order (Peano-Hilbert order of
        the cells) and the sorted 48-bit octree keys."""
        si = np.zeros(max(self.n, 1), np.int32)
        wi = np.zeros(max(self.n, 1), np.int32)
        ks = np.zeros(max(self.n, 1), np.uint64)
        nl = ctypes.c_int(0)
        _check(self.L.cosmosim_ext_debug_orders(self.h, ptr(si), ptr(wi), ptr(ks), self.n,
                                                ctypes.byref(nl)), "debug_orders")
        k = nl.value
        return si[:k].copy(), wi[:k].copy(), ks[:k].copy()

    # ------------------------------------------------------------ adjoint
    def force_vjp(self, w):
        """(d a / d x)^T w at the current positions (tree and decisions frozen)."""
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.n, 3)
        out = np.zeros((self.n, 3))
        _check(self.L.cosmosim_ext_force_vjp(self.h, ptr(w), ptr(out)), "force_vjp")
        return out

    def forces_frozen(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        out = np.zeros((self.n, 3))
        _check(self.L.cosmosim_ext_forces_frozen(self.h, ptr(pos), ptr(out)), "forces_frozen")
        return out

    def direct_vjp(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.n, 3)
        out = np.zeros((self.n, 3))
        _check(self.L.cosmosim_ext_direct_vjp(self.h, ptr(w), ptr(out)), "direct_vjp")
        return out

    def record(self, substeps, checkpoint_every=10, keep_trees=False):
        _check(self.L.cosmosim_ext_record(self.h, substeps, checkpoint_every, int(keep_trees)),
               "record")

    def backward(self, gpos, gvel=None, mode=0):
        """(dL/dx0, dL/dv0, stats) from dL/dx_k (and dL/dv_k) of the recorded run."""
        gp = np.ascontiguousarray(gpos, dtype=np.float64).reshape(self.n, 3)
        gv = None if gvel is None else np.ascontiguousarray(gvel, dtype=np.float64).reshape(self.n, 3)
        xb, vb, st = np.zeros((self.n, 3)), np.zeros((self.n, 3)), np.zeros(2)
        _check(self.L.cosmosim_ext_backward(self.h, ptr(gp), ptr(gv), mode, ptr(xb), ptr(vb),
                                           ptr(st)), "backward")
        return xb, vb, st

    def replay(self, pos0, vel0):
        p0 = np.ascontiguousarray(pos0, dtype=np.float64).reshape(self.n, 3)
        v0 = np.ascontiguousarray(vel0, dtype=np.float64).reshape(self.n, 3)
        p, v = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        _check(self.L.cosmosim_ext_replay(self.h, ptr(p0), ptr(v0), ptr(p), ptr(v)), "replay")
        return p, v

