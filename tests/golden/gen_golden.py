"""Write the golden fixtures in tests/golden/ from the reference's own C code.

Runs only where the reference sources were compiled into
oracle/_ref/libcosmosim_ref.so (oracle/Makefile, target `ref`).  Every fixture
holds inputs and the reference's outputs for them:

  kat.npz        the hand-built cases of tests/test_physics.c (single body, two
                 bodies, 100-body lattice, 8 bodies, dead body, 2-body orbit over
                 2000 steps, 4-body momentum over 500 steps) plus n=1, all-dead,
                 coincident bodies (depth-64 merge) and two-body chains
  disk4096.npz   seeded 4096-body disk (every 37th body dead): reference pool and
                 accelerations at theta 0, 0.4, 0.5, 0.7 (eps 0.01), neighbour queries
  traj1024.npz   1024-body merger, 10 KDK substeps (theta 0.5, eps 0.01, dt 1e-3)

    python tests/golden/gen_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from cosmosim_amd._abi import make_bodies, BODY_DTYPE  # noqa: E402
from oracle.cpu_ref import RefBuild  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
HOT = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "mass"]


def hot(b):
    return np.stack([b[f] for f in HOT], axis=1)


def kat_inputs():
    """Inputs restated from tests/test_physics.c (line numbers there)."""
    c = {}
    c["single"] = make_bodies([[3.0, 4.0, 5.0]], mass=[7.0])                      # :87-104
    c["two"] = make_bodies([[1, 0, 0], [-1, 0, 0]], mass=[1.0, 3.0])              # :106-124
    i = np.arange(100)
    lat = np.stack([(i * 17 % 97) - 48.0, (i * 31 % 89) - 44.0, (i * 13 % 73) - 36.0], 1)
    c["lattice"] = make_bodies(lat, mass=1.0 + (i % 5))                            # :126-148
    c["two_x5"] = make_bodies([[0, 0, 0], [5, 0, 0]], mass=[1.0, 1.0])             # :152-175
    c["newton3"] = make_bodies([[0, 0, 0], [3, 4, 0]], mass=[2.0, 5.0])            # :177-199
    pos8 = [[1, 2, 3], [-4, 5, -1], [7, -3, 2], [0, 0, 0], [-2, -6, 4], [3, 1, -5],
            [-1, 8, 0], [6, -2, -3]]
    c["kat8"] = make_bodies(pos8, mass=[1, 3, 2, 5, 1, 4, 2, 3])                   # :234-270
    c["dead3"] = make_bodies([[0, 0, 0], [5, 0, 0], [-5, 0, 0]], mass=[10.0, 0.0, 3.0])  # :396-421
    c["n1"] = make_bodies([[0.25, -1.5, 2.0]], mass=[2.0])
    c["alldead"] = make_bodies([[1, 2, 3], [4, 5, 6]], mass=[0.0, -1.0])
    c["coincident"] = make_bodies([[1, 1, 1], [1, 1, 1], [1, 1, 1], [2, 2, 2], [1, 1, 1 + 1e-12]],
                                  mass=[1, 2, 3, 4, 5])
    c["chain"] = make_bodies([[0, 0, 0], [1e-7, 2e-7, -3e-7], [10, 10, 10]], mass=[1, 1, 1])
    return c


def main():
    if not RefBuild.available():
        raise SystemExit("oracle/_ref/libcosmosim_ref.so is missing: run `make -C oracle ref`")
    ref = RefBuild()
    out = {}
    for name, b in kat_inputs().items():
        out[f"{name}_in"] = hot(b)
        out[f"{name}_pool"] = ref.build(b)
        for th, eps in [(0.0, 0.01), (0.5, 0.01), (0.0, 0.1), (0.0, 1e-4)]:
            bb = b.copy()
            pool = ref.build(bb)
            ref.forces(pool, bb, 1.0, eps, th)
            out[f"{name}_acc_t{th}_e{eps}"] = hot(bb)[:, 6:9]
    # 2-body circular orbit, 2000 steps (:274-304)
    M, r = 100.0, 5.0
    b = make_bodies([[0, 0, 0], [r, 0, 0]], vel=[[0, 0, 0], [0, np.sqrt(M / r), 0]], mass=[M, 1.0])
    out["orbit_in"] = hot(b)
    ref.init_acc(b, 1.0, 0.01, 0.0)
    for _ in range(2000):
        ref.step(b, 1e-3, 1.0, 0.01, 0.0)
    out["orbit_out"] = hot(b)
    # 4-body momentum, 500 steps (:306-343)
    b = make_bodies([[1, 0, 0], [-1, 0, 0], [0, 2, 1], [0, -2, -1]],
                    vel=[[0, 1, 0], [0, -1, 0], [-0.5, 0, 0.3], [0.5, 0, -0.3]],
                    mass=[2.0, 2.0, 3.0, 1.0])
    out["mom4_in"] = hot(b)
    ref.init_acc(b, 1.0, 0.1, 0.0)
    for _ in range(500):
        ref.step(b, 2e-3, 1.0, 0.1, 0.0)
    out["mom4_out"] = hot(b)
    np.savez_compressed(os.path.join(OUT, "kat.npz"), **out)

    # seeded disk (inputs stored, so the fixture does not depend on the generator)
    import cosmosim_amd as ca
    d = ca.generate(ca.IC_DISK, 4096, 11)
    d["mass"][::37] = 0.0
    dk = {"bodies": hot(d), "pool": ref.build(d)}
    for th in [0.0, 0.4, 0.5, 0.7]:
        bb = d.copy()
        ref.forces(dk["pool"], bb, 1.0, 0.01, th)
        dk[f"acc_t{th}"] = hot(bb)[:, 6:9]
    rng = np.random.default_rng(7)
    centres = rng.normal(size=(24, 3)) * np.array([1.5, 1.5, 0.1])
    res, offs = [], [0]
    for c in centres:
        nb = ref.find_neighbors(dk["pool"], d, c, 0.25, 128)
        res.append(nb)
        offs.append(offs[-1] + len(nb))
    dk["nb_centres"] = centres
    dk["nb_radius"] = np.array(0.25)
    dk["nb_cap"] = np.array(128)
    dk["nb_result"] = np.concatenate(res).astype(np.int32)
    dk["nb_offsets"] = np.array(offs, dtype=np.int64)
    np.savez_compressed(os.path.join(OUT, "disk4096.npz"), **dk)

    m = ca.generate(ca.IC_MERGER, 1024, 4)
    tr = {"bodies": hot(m)}
    ref.init_acc(m, 1.0, 0.01, 0.5)
    tr["init"] = hot(m)
    for _ in range(10):
        ref.step(m, 1e-3, 1.0, 0.01, 0.5)
    tr["final"] = hot(m)
    np.savez_compressed(os.path.join(OUT, "traj1024.npz"), **tr)
    for f in ["kat.npz", "disk4096.npz", "traj1024.npz"]:
        print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
