"""Write tests/golden/sph.npz from the reference's own C code (src/sph.c).

Runs only where the reference sources were compiled into
oracle/_ref/libcosmosim_ref.so (oracle/Makefile, target `ref`).  Inputs and the
reference's outputs of sph_compute_density, sph_compute_forces (accelerations
zeroed first, as the reference's SPH tests do) and sph_apply_cooling:

  kat_*     the inputs of the reference's SPH tests (tests/test_physics.c:838-1043)
  mix_*     1,500 bodies from the seeded merger generator, about two thirds made
            gas with scattered smoothing lengths and internal energies, every
            41st body dead, five coincident gas pairs

    python tests/golden/gen_golden_sph.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cosmosim_amd._abi import make_bodies  # noqa: E402
from cosmosim_amd import generate, IC_MERGER  # noqa: E402
from oracle.cpu_ref import RefBuild  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
GAS = 1


def kat_inputs():
    """The SPH test inputs of tests/test_physics.c (line numbers there)."""
    c = {}
    i = np.arange(200)                                                     # :838-877
    pos = np.stack([(i % 10) / 9.0 * 10.0 - 5.0, ((i // 10) % 10) / 9.0 * 10.0 - 5.0,
                    (i // 100) / 1.0 * 2.0 - 1.0], 1)
    c["cube200"] = make_bodies(pos, mass=1.0)
    g = np.array([[-1.5 + ix, -1.5 + iy, -1.5 + iz] for ix in range(4) for iy in range(4)
                  for iz in range(4)], dtype=float)                        # :879-926
    b = make_bodies(g, mass=1.0)
    b["type"], b["internal_energy"], b["smoothing_h"] = GAS, 1.0, 1.5
    c["grid64"] = b
    x = np.concatenate([-3.0 + np.arange(30) * 0.1, np.arange(10) * 0.3])  # :928-980
    b = make_bodies(np.stack([x, 0 * x, 0 * x], 1), mass=1.0)
    b["type"], b["internal_energy"], b["smoothing_h"] = GAS, 1.0, 0.5
    c["interface40"] = b
    x = np.arange(20) * 0.5                                                # :982-1016
    b = make_bodies(np.stack([x, 0 * x, 0 * x], 1), mass=1.0)
    b["type"] = np.where(np.arange(20) % 2 == 0, GAS, 0)
    b["internal_energy"], b["smoothing_h"] = 1.0, 2.0
    c["mixed20"] = b
    b = make_bodies(np.stack([np.arange(5.0), np.zeros(5), np.zeros(5)], 1), mass=1.0)  # :1020-1043
    b["type"], b["internal_energy"], b["density"] = GAS, 100.0, 1.0
    c["hot5"] = b
    return c


def mix_input():
    rng = np.random.default_rng(11)
    b = generate(IC_MERGER, 1500, 6)
    gas = rng.uniform(size=1500) < 0.67
    b["type"] = np.where(gas, GAS, b["type"])
    b["smoothing_h"] = np.where(gas, rng.uniform(0.05, 0.8, 1500), 0.0)
    b["internal_energy"] = np.where(gas, rng.uniform(0.0, 2.0, 1500), 0.0)
    b["internal_energy"][::97] = 0.0          # the 1e-15 floor
    b["mass"][::41] = 0.0
    gi = np.nonzero(gas & (b["mass"] > 0))[0]
    for k in range(5):                         # coincident gas pairs (r < 1e-15 skipped)
        a, c = gi[2 * k], gi[2 * k + 1]
        for f in ("x", "y", "z"):
            b[f][c] = b[f][a]
    return b


def run(R, b):
    """Density, forces (a zeroed first) and cooling; returns the three states."""
    pool = R.build(b)
    d = b.copy()
    R.sph_density(d, pool)
    f = d.copy()
    f["ax"] = f["ay"] = f["az"] = 0.0
    R.sph_forces(f, pool)
    c = f.copy()
    R.sph_cooling(c, 1e-2)
    return d, f, c


def main():
    if not RefBuild.available():
        raise SystemExit("oracle/_ref/libcosmosim_ref.so is missing: run `make -C oracle ref`")
    R = RefBuild()
    out = {}
    cases = kat_inputs()
    cases["mix"] = mix_input()
    for name, b in cases.items():
        d, f, c = run(R, b)
        out[f"{name}_in"] = b
        out[f"{name}_density"], out[f"{name}_forces"], out[f"{name}_cooling"] = d, f, c
    np.savez_compressed(os.path.join(OUT, "sph.npz"), **out)
    print("wrote", os.path.join(OUT, "sph.npz"))


if __name__ == "__main__":
    main()
