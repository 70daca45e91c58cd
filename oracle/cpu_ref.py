"""ctypes access to the CPU oracle.  TEST INFRASTRUCTURE ONLY.

Two implementations with one Python interface:

* ``CpuRef``   - oracle/libcpuref.so, the clean-room restatement in cpu_ref.c
                 (travels with the repository; built by oracle/Makefile);
* ``RefBuild`` - oracle/_ref/libcosmosim_ref.so, the reference's own C sources
                 (src/octree.c, src/integrator.c, ...) compiled in place, present
                 only where the reference tree was available at build time.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker or the timed CPU baseline, never as the product.
"""
import ctypes
import os
import subprocess

import numpy as np

from cosmosim_amd._abi import BODY_DTYPE, NODE_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
CPUREF_SO = os.path.join(HERE, "libcpuref.so")
REF_SO = os.path.join(HERE, "_ref", "libcosmosim_ref.so")

class QuasarConfig(ctypes.Structure):
    """src/quasar.h:6-22 (the defaults of quasar_default_config, src/quasar.c:38-56)."""
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


_VP = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int


def build(quiet=True):
    """Compile oracle/libcpuref.so (and _ref when the reference is present)."""
    out = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _pool_for(n):
    # 8n as the reference's callers allocate, plus room for depth-64 chains of
    # near-coincident bodies (the reference itself has no bound check)
    return np.zeros(8 * max(n, 1) + 64 * min(n, 1024) + 16, dtype=NODE_DTYPE)


class _Base:
    def build(self, bodies):
        """Reference-layout tree of the live bodies; returns the used pool slice."""
        pool = _pool_for(len(bodies))
        used = _I(0)
        self._build(_p(pool), ctypes.byref(used), _p(bodies), len(bodies))
        return pool[: used.value].copy()

    def init_acc(self, bodies, G, softening, theta):
        pool = _pool_for(len(bodies))
        self._init(_p(bodies), len(bodies), _D(G), _D(softening), _D(theta), _p(pool))

    def step(self, bodies, dt, G, softening, theta, pool=None):
        if pool is None:
            pool = _pool_for(len(bodies))
        self._step_impl(bodies, dt, G, softening, theta, pool)

    def find_neighbors(self, pool, bodies, c, radius, cap):
        buf = np.zeros(max(cap, 1), dtype=np.int32)
        k = self._nb(_p(pool), 0, _p(bodies), len(bodies), _D(c[0]), _D(c[1]), _D(c[2]),
                     _D(radius), _p(buf), cap)
        return buf[:k].copy()

    # SPH passes over a body array (in place) with the tree `pool` (root 0)
    def sph_density(self, bodies, pool):
        self._sph_density(_p(bodies), len(bodies), _p(pool))

    def sph_forces(self, bodies, pool):
        self._sph_forces(_p(bodies), len(bodies), _p(pool))

    def sph_cooling(self, bodies, dt):
        self._sph_cooling(_p(bodies), len(bodies), _D(dt))


class CpuRef(_Base):
    kind = "port"

    def __init__(self, path=CPUREF_SO):
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        self.lib = L
        L.cref_build.argtypes = [_VP, _VP, _VP, _I]
        L.cref_forces.argtypes = [_VP, _I, _VP, _I, _D, _D, _D, _VP]
        L.cref_list_walk.argtypes = [_VP, _I, _VP, _I, _VP, _I, _D, _D, _D, _VP, _VP, _VP]
        L.cref_init_acc.argtypes = [_VP, _I, _D, _D, _D, _VP]
        L.cref_step.argtypes = [_VP, _I, _D, _D, _D, _D, _VP]
        L.cref_kick_drift.argtypes = [_VP, _I, _D]
        L.cref_kick.argtypes = [_VP, _I, _D]
        L.cref_direct.argtypes = [_VP, _I, _VP, _I, _D, _D, _VP]
        L.cref_energy.argtypes = [_VP, _I, _D, _D, _VP]
        L.cref_find_neighbors.argtypes = [_VP, _I, _VP, _I, _D, _D, _D, _D, _VP, _I]
        L.cref_find_neighbors.restype = _I
        L.cref_canonical.argtypes = [_VP, _I, _VP, _VP, _I]
        L.cref_canonical.restype = _I
        for f in ("cref_sph_density", "cref_sph_forces"):
            getattr(L, f).argtypes = [_VP, _I, _VP]
        L.cref_sph_cooling.argtypes = [_VP, _I, _D]
        L.cref_snapshot.argtypes = [_VP, _I, _I, _VP, _I]
        L.cref_snapshot.restype = _I
        L.cref_quasar_step.argtypes = [_VP, _VP, _I, _VP, _D, _VP, _VP]
        L.cref_decay_jets.argtypes = [_VP, _VP, _I, _VP, _D, _VP]
        L.cref_quasar_compact.argtypes = [_VP, _I]
        L.cref_quasar_compact.restype = _I
        self._build = L.cref_build
        self._init = L.cref_init_acc
        self._nb = L.cref_find_neighbors
        self._sph_density = L.cref_sph_density
        self._sph_forces = L.cref_sph_forces
        self._sph_cooling = L.cref_sph_cooling

    def forces(self, pool, bodies, G, softening, theta, root=0):
        """Zero a, walk every live body; returns per-body interaction counts."""
        counts = np.zeros(len(bodies), dtype=np.int32)
        self.lib.cref_forces(_p(pool), root, _p(bodies), len(bodies), _D(G), _D(softening),
                             _D(theta), _p(counts))
        return counts

    def list_walk(self, pool, bodies, order, G, softening, theta, root=0):
        """The core's interaction-list walk restated (cref_list_walk): groups of 64
        consecutive bodies of `order`.  Returns (acc (n, 3), per-body list counts,
        worst relative mass-coverage error of a group's list, number of listed
        internal nodes that fail the per-body criterion)."""
        order = np.ascontiguousarray(order, dtype=np.int32)
        acc = np.zeros((len(bodies), 3))
        counts = np.zeros(len(bodies), dtype=np.int32)
        checks = np.zeros(2)
        self.lib.cref_list_walk(_p(pool), root, _p(bodies), len(bodies), _p(order), len(order),
                                _D(G), _D(softening), _D(theta), _p(acc), _p(counts), _p(checks))
        return acc, counts, float(checks[0]), int(checks[1])

    def _step_impl(self, bodies, dt, G, softening, theta, pool):
        self.lib.cref_step(_p(bodies), len(bodies), _D(dt), _D(G), _D(softening), _D(theta),
                           _p(pool))

    def kick_drift(self, bodies, dt):
        self.lib.cref_kick_drift(_p(bodies), len(bodies), _D(dt))

    def kick(self, bodies, dt):
        self.lib.cref_kick(_p(bodies), len(bodies), _D(dt))

    def direct(self, bodies, idx, G, softening):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        acc = np.zeros((len(idx), 3))
        self.lib.cref_direct(_p(bodies), len(bodies), _p(idx), len(idx), _D(G), _D(softening),
                             _p(acc))
        return acc

    def energy(self, bodies, G, softening):
        out = np.zeros(2)
        self.lib.cref_energy(_p(bodies), len(bodies), _D(G), _D(softening), _p(out))
        return float(out[0]), float(out[1])

    def quasar_step(self, bodies, n, cfg, dt, state):
        """One quasar step on bodies[:n] (capacity len(bodies)); state = [rng, phase]
        carried between calls (the reference's statics start at rng 1, phase 0).
        Returns the new n; cfg (QuasarConfig) is updated in place."""
        nn = _I(n)
        rng = ctypes.c_uint64(state[0])
        ph = _I(state[1])
        self.lib.cref_quasar_step(_p(bodies), ctypes.byref(nn), len(bodies), ctypes.byref(cfg),
                                  _D(dt), ctypes.byref(rng), ctypes.byref(ph))
        state[0], state[1] = rng.value, ph.value
        return nn.value

    def decay_jets(self, bodies, n, cfg, dt, state):
        nn = _I(n)
        rng = ctypes.c_uint64(state[0])
        self.lib.cref_decay_jets(_p(bodies), ctypes.byref(nn), len(bodies), ctypes.byref(cfg),
                                 _D(dt), ctypes.byref(rng))
        state[0] = rng.value
        return nn.value

    def quasar_compact(self, bodies, n):
        return self.lib.cref_quasar_compact(_p(bodies), n)

    def snapshot(self, bodies, fmt, max_out=None):
        """Float snapshot of the live bodies (fmt 0: 16 floats, 1: 10 floats)."""
        m = len(bodies) if max_out is None else max_out
        out = np.zeros((max(m, 1), 16 if fmt == 0 else 10), dtype=np.float32)
        k = self.lib.cref_snapshot(_p(bodies), len(bodies), fmt, _p(out), m)
        return out[:k].copy()

    def canonical(self, pool, root=0):
        """Pool renumbered in depth-first pre-order, children in slot order."""
        out = np.zeros(len(pool), dtype=NODE_DTYPE)
        m = np.zeros(len(pool), dtype=np.int32)
        k = self.lib.cref_canonical(_p(pool), root, _p(out), _p(m), len(pool))
        return out[:k].copy()


class RefBuild(_Base):
    """The reference's compiled sources (pins CpuRef; present only in the build container)."""

    kind = "reference"

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def __init__(self, path=REF_SO):
        L = ctypes.CDLL(path)
        self.lib = L
        L.octree_build.argtypes = [_VP, _VP, _VP, _I]
        L.octree_compute_forces.argtypes = [_VP, _I, _VP, _I, _D, _D, _D]
        L.integrator_init_accelerations.argtypes = [_VP, _I, _D, _D, _D, _VP]
        L.integrator_step.argtypes = [_VP, _I, _D, _D, _D, _D, _VP, _I]
        L.octree_find_neighbors.argtypes = [_VP, _I, _VP, _I, _D, _D, _D, _D, _VP, _I]
        L.octree_find_neighbors.restype = _I
        for f in ("sph_compute_density", "sph_compute_forces"):
            getattr(L, f).argtypes = [_VP, _I, _VP]
        L.sph_apply_cooling.argtypes = [_VP, _I, _D]
        self._build = L.octree_build
        self._init = L.integrator_init_accelerations
        self._nb = L.octree_find_neighbors
        self._sph_density = L.sph_compute_density
        self._sph_forces = L.sph_compute_forces
        self._sph_cooling = L.sph_apply_cooling

    def forces(self, pool, bodies, G, softening, theta, root=0):
        self.lib.octree_compute_forces(_p(pool), root, _p(bodies), len(bodies), _D(G),
                                       _D(softening), _D(theta))
        return None

    def _step_impl(self, bodies, dt, G, softening, theta, pool):
        self.lib.integrator_step(_p(bodies), len(bodies), _D(dt), _D(G), _D(softening),
                                 _D(theta), _p(pool), 0)


class RefQuasar:
    """The reference's quasar_step / quasar_compact from a private copy of
    oracle/_ref's library: its random stream and knot phase are file-scope statics,
    so each instance starts from the reference's initial state (rng 1, phase 0)."""

    def __init__(self):
        import shutil
        import tempfile
        self._dir = tempfile.mkdtemp(prefix="refq")
        path = os.path.join(self._dir, "libref_copy.so")
        shutil.copy(REF_SO, path)
        L = ctypes.CDLL(path)
        L.quasar_step.argtypes = [_VP, _VP, _I, _VP, _D]
        L.decay_jets_wrapper.argtypes = [_VP, _VP, _I, _VP, _D]
        L.quasar_compact.argtypes = [_VP, _I]
        L.quasar_compact.restype = _I
        self.lib = L

    def quasar_step(self, bodies, n, cfg, dt):
        nn = _I(n)
        self.lib.quasar_step(_p(bodies), ctypes.byref(nn), len(bodies), ctypes.byref(cfg), _D(dt))
        return nn.value

    def decay_jets(self, bodies, n, cfg, dt):
        nn = _I(n)
        self.lib.decay_jets_wrapper(_p(bodies), ctypes.byref(nn), len(bodies), ctypes.byref(cfg),
                                    _D(dt))
        return nn.value

    def quasar_compact(self, bodies, n):
        return self.lib.quasar_compact(_p(bodies), n)


def bodies_copy(b):
    return np.array(b, dtype=BODY_DTYPE, copy=True)
