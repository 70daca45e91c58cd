"""The C-ABI library and the host logic - CPU only (no kernel is launched).

* libcosmosim.so loads and exports every function declared in include/*.h;
* the C records have the reference's layout (a C program compiled against
  include/*.h reports sizeof/offsetof, compared with the numpy/ctypes mirrors);
* the seeded initial-condition generator is deterministic and follows its recipe;
* without a HIP device every entry point fails loudly (no CPU fallback).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import cosmosim_amd as ca
from cosmosim_amd._abi import BODY_DTYPE, NODE_DTYPE, CosmosimConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = ["body.h", "octree.h", "integrator.h", "cosmosim_api.h", "cosmosim_ext.h", "csg.h"]


def _declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(INC, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:COSMOSIM_API\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(",
                             text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "sizeof", "offsetof", "defined"):
                names.add(name)
    return names


def test_library_exports_every_declared_symbol():
    declared = _declared_functions()
    assert len(declared) >= 50
    L = ca.lib()
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert declared == set(ca.EXPORTED_SYMBOLS), declared ^ set(ca.EXPORTED_SYMBOLS)
    out = subprocess.run(["nm", "-D", "--defined-only", ca.library_path], capture_output=True,
                         text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert declared <= exported


def test_layouts_match_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r"""
#include <stdio.h>
#include <stddef.h>
#include "body.h"
#include "octree.h"
#include "cosmosim_api.h"
#include "cosmosim_ext.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(Body), sizeof(OctreeNode), sizeof(CosmosimConfig),
         sizeof(CosmosimExtParams), sizeof(CosmosimExtStats));
  printf("%zu %zu %zu %zu\n", offsetof(Body, type), offsetof(Body, smoothing_h),
         offsetof(OctreeNode, children), offsetof(CosmosimConfig, substeps));
  return 0; }
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I" + INC, str(src), "-o", str(exe)], check=True)
    a, b = subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")[:2]
    assert [int(v) for v in a.split()] == [168, 104, 72, ctypes.sizeof(ca.ExtParams),
                                            ctypes.sizeof(ca.ExtStats)]
    assert [int(v) for v in b.split()] == [BODY_DTYPE.fields["type"][1],
                                           BODY_DTYPE.fields["smoothing_h"][1],
                                           NODE_DTYPE.fields["children"][1],
                                           CosmosimConfig.substeps.offset]


def test_default_config_matches_reference():
    cfg = ca.lib().cosmosim_default_config()   # src/cosmosim_api.c:30-45
    assert (cfg.n_bodies, cfg.merger, cfg.quasar, cfg.substeps) == (20000, 0, 0, 2)
    assert (cfg.dt, cfg.theta, cfg.smbh_mass_frac) == (0.005, 0.5, 0.05)
    assert (cfg.accretion_radius, cfg.jet_speed, cfg.feedback_strength) == (6.0, 15.0, 0.3)


def test_generator_disk_recipe():
    n = 20001
    b = ca.generate(ca.IC_DISK, n, 1)
    assert np.array_equal(b.view(np.uint8), ca.generate(ca.IC_DISK, n, 1).view(np.uint8))
    assert not np.array_equal(b["x"], ca.generate(ca.IC_DISK, n, 2)["x"])
    assert b["mass"][0] == 0.05 and b["type"][0] == ca._abi.BODY_SMBH
    assert b["x"][0] == b["y"][0] == b["z"][0] == 0.0
    assert np.allclose(b["mass"][1:], 1.0 / (n - 1), rtol=0, atol=0)
    R = np.hypot(b["x"][1:], b["y"][1:])
    assert R.max() <= 5.0 and 1.6 < R.mean() < 2.2            # <R> of R e^-R on [0,5] ~ 1.87
    assert 0.09 < np.std(b["z"][1:]) < 0.11
    # rotation in +phi at about the circular speed
    lz = b["x"][1:] * b["vy"][1:] - b["y"][1:] * b["vx"][1:]
    assert np.mean(lz > 0) > 0.99


def test_generator_merger_recipe_and_orbit_coefficients():
    n = 10000
    b = ca.generate(ca.IC_MERGER, n, 2)
    smbh = np.nonzero(b["type"] == ca._abi.BODY_SMBH)[0]
    assert len(smbh) == 2 and smbh[0] == 0
    assert np.isclose(b["mass"][smbh[0]], 0.05) and np.isclose(b["mass"][smbh[1]], 0.06)
    stars = b["mass"][b["type"] == 0]
    assert np.allclose(stars, 2.2 / (n - 2))
    assert abs(b["mass"].sum() - 2.31) < 1e-12
    # centre-of-mass frame: total momentum ~ 0 (dispersion noise only)
    p = np.array([np.sum(b["mass"] * b[f]) for f in ["vx", "vy", "vz"]])
    assert np.linalg.norm(p) < 0.02
    coef = ca.orbit_coefficients(n)
    orb = ca.default_orbit()
    d = np.array([0.5, -0.25, 1.0, 0.01, 0.02, -0.03])
    b2 = ca.generate(ca.IC_MERGER, n, 2, orbit=orb + d)
    for k, f in enumerate(["x", "y", "z", "vx", "vy", "vz"]):
        assert np.allclose(b2[f] - b[f], coef * d[k], rtol=0, atol=1e-12)
    assert np.allclose(np.hypot(orb[0], orb[1]), 30.0) and orb[3] == -0.3


def test_error_strings():
    L = ca.lib()
    assert L.cosmosim_ext_error_string(0) == b"ok"
    assert b"device" in L.cosmosim_ext_error_string(-2)
    assert L.cosmosim_ext_generate(7, 10, 1, None, None) == -1


def _have_device():
    k = ctypes.c_int(0)
    return ca.lib().csg_device_count(ctypes.byref(k)) == 0 and k.value > 0


@pytest.mark.skipif(_have_device(), reason="a HIP device is present")
def test_without_device_every_path_fails_loudly():
    b = ca.make_bodies(np.random.default_rng(0).normal(size=(16, 3)))
    with pytest.raises(ca.CosmosimError):
        ca.octree_build(b)
    with pytest.raises(ca.CosmosimError):
        ca.integrator_step(b, 1e-3, 1.0, 0.01, 0.5)
    with pytest.raises(ca.CosmosimError):
        ca.Simulation(b)
    assert not ca.lib().cosmosim_create(ca.lib().cosmosim_default_config())
    # the handle API tolerates NULL (src/cosmosim_api.c:117,148,155,162,175,182)
    L = ca.lib()
    L.cosmosim_step(None)
    L.cosmosim_destroy(None)
    assert L.cosmosim_get_bodies(None) is None and L.cosmosim_get_count(None) == 0
    assert L.cosmosim_get_active_count(None) == 0 and L.cosmosim_get_sim_time(None) == 0.0
