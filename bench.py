"""Benchmark: substeps/s of the 1M-body merger (FP32 walk vs FP64 walk) on one GPU.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--thetas 0.4,0.5,0.7]

A step is one full KDK substep on the device (kick+drift, tree build, walk,
kick) of the seeded 1,000,000-body two-disk merger (seed 3, eps 0.005,
dt 5e-4), with every input resident in HBM before the clock starts.  The
headline value is the FP32 walk on wave-shared interaction lists at theta 0.5
(each wavefront's 64 bodies share one list of the cells that pass the opening
test at their bounding box, which is at least the reference test for each of
them); the FP32 walk with the reference's per-body test and the FP64 walks
(strict reference arithmetic and refined-rsqrt FP64) are timed alongside at
each theta, every walk with its error against the FP64 direct sum.  Prints one
JSON line.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "substeps/s @1M-body merger: FP32 walk vs FP64, at RMS accel err; adjoint step/s"
N_BODIES = 1_000_000
SEED = 3
EPS = 0.005
DT = 5e-4
FLOP_PER_INTERACTION = 20       # 3 sub, 5 for d^2, +eps^2, sqrt, mul, div, G*M, 6 scale-add
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
SUBSET = 4096


def rel_err(a, b):
    d = np.linalg.norm(a - b, axis=1)
    s = np.linalg.norm(b, axis=1)
    s[s == 0] = 1.0
    e = d / s
    return float(np.sqrt(np.mean(e ** 2))), float(e.max())


def timed_run(ca, bodies, theta, precision, steps, warmup, energy=False):
    sim = ca.Simulation(bodies, softening=EPS, theta=theta, dt=DT, precision=precision)
    sim.advance(warmup)
    e0 = sim.energy() if energy else None
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.advance(steps)                       # returns after the stream has drained
    t1 = time.perf_counter()
    st = sim.stats()
    out = {
        "substeps_per_s": steps / (t1 - t0),
        "ms_per_step": 1e3 * (t1 - t0) / steps,
        "walk_ms": st.ms_walk / max(st.timed_walks, 1),
        "build_ms": st.ms_build / steps,
        "kdk_ms": st.ms_kdk / steps,
        "interactions": int(st.interactions),
        "nodes": int(st.n_nodes),
        "max_depth": int(st.max_depth),
    }
    out["interactions_per_s"] = out["interactions"] * out["substeps_per_s"]
    if energy:
        e1 = sim.energy()
        E0, E1 = e0[0] + e0[1], e1[0] + e1[1]
        out["energy_drift"] = abs((E1 - E0) / E0)
        out["momentum_drift"] = float(np.linalg.norm(e1[2] - e0[2]))
    return sim, out


def accuracy(ca, sim, idx):
    """Errors on one state: the per-body FP32 walk vs FP64 on the same tree
    (precision), FP64 vs FP64 direct sum (method), each FP32 walk vs the direct
    sum, and the list walk vs the per-body FP64 walk (the two opening rules)."""
    sim.set_params(precision="fp64_strict")
    sim.compute_forces()
    a64 = sim.accelerations()
    sim.set_params(precision="fp32")
    sim.compute_forces()
    a32 = sim.accelerations()
    sim.set_params(precision="fp32_list")
    sim.compute_forces()
    al = sim.accelerations()
    direct = sim.direct_forces(idx)
    r = {}
    r["fp32_vs_fp64_rms"], r["fp32_vs_fp64_max"] = rel_err(a32[idx], a64[idx])
    r["fp32_vs_fp64_rms_all"], r["fp32_vs_fp64_max_all"] = rel_err(a32, a64)
    r["fp64_vs_direct_rms"], r["fp64_vs_direct_max"] = rel_err(a64[idx], direct)
    r["fp32_vs_direct_rms"], r["fp32_vs_direct_max"] = rel_err(a32[idx], direct)
    r["fp32_list_vs_direct_rms"], r["fp32_list_vs_direct_max"] = rel_err(al[idx], direct)
    r["fp32_list_vs_fp64_rms_all"], r["fp32_list_vs_fp64_max_all"] = rel_err(al, a64)
    return r


def per_step_accuracy(ca, bodies, idx, theta, steps, precision="fp32_list"):
    """FP32 substeps of the 1M merger (untimed): after every substep, the FP32
    walk's accelerations of the 4096-body subset against the FP64 direct sum on
    the same state (error (c)); errors (a) and (b) are taken at the first and
    last state by accuracy()."""
    sim = ca.Simulation(bodies, softening=EPS, theta=theta, dt=DT, precision=precision)
    rms, mx = [], []
    for _ in range(steps):
        sim.advance(1)
        a = sim.accelerations()
        r, m = rel_err(a[idx], sim.direct_forces(idx))
        rms.append(r)
        mx.append(m)
    sim.close()
    return {"steps": steps, "fp32_vs_direct_rms_mean": float(np.mean(rms)),
            "fp32_vs_direct_rms_max": float(np.max(rms)),
            "fp32_vs_direct_max_max": float(np.max(mx))}


def config_line(ca, recipe, n, seed, theta, eps, dt, steps, precision, energy=True):
    """Substeps/s, interactions/s and the walk's roofline fraction of one of the
    smaller configurations, with its energy and momentum drift over the run (the
    energy's potential is a direct sum: energy=False skips it at large n)."""
    sim = ca.Simulation(ca.generate(recipe, n, seed), softening=eps, theta=theta, dt=dt,
                        precision=precision)
    ke0, pe0, p0 = sim.energy() if energy else (0.0, 0.0, np.zeros(3))
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.advance(steps)
    t1 = time.perf_counter()
    st = sim.stats()
    ke1, pe1, p1 = sim.energy() if energy else (0.0, 0.0, np.zeros(3))
    sim.close()
    walk_ms = st.ms_walk / max(st.timed_walks, 1)
    peak = FP32_PEAK_TFLOPS if precision.startswith("fp32") else FP64_PEAK_TFLOPS
    tflops = st.interactions * FLOP_PER_INTERACTION / (walk_ms * 1e-3) / 1e12
    return {"bodies": n, "substeps": steps, "precision": precision,
            "substeps_per_s": steps / (t1 - t0), "walk_ms": walk_ms,
            "build_ms": st.ms_build / steps, "kdk_ms": st.ms_kdk / steps,
            "interactions": int(st.interactions),
            "interactions_per_s": st.interactions / (walk_ms * 1e-3),
            "walk_tflops": tflops, "walk_valu_frac": tflops / peak,
            **({"energy_drift": abs((ke1 + pe1 - ke0 - pe0) / (ke0 + pe0)),
                "momentum_drift": float(np.linalg.norm(p1 - p0))} if energy else {})}


def cpu_baseline(ca, bodies, theta):
    """The oracle's restatement of the reference (oracle/libcpuref.so: serial
    insertion build, the walk under the reference's OpenMP schedule, serial KDK)
    on the host cores: median of 3 KDK substeps."""
    from oracle.cpu_ref import CpuRef
    impl = CpuRef()
    b = bodies.copy()
    impl.init_acc(b, 1.0, EPS, theta)
    pool = np.zeros(8 * len(b) + 16, dtype=ca.NODE_DTYPE)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        impl.step(b, DT, 1.0, EPS, theta, pool)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / med, "unit": "substeps/s", "cores": threads, "kind": impl.kind,
        "sample": (f"3 KDK substeps (median) of the same 1M-body merger at theta {theta}, "
                   f"eps {EPS}: serial build, OpenMP walk ({threads} threads), serial KDK"),
        "host_cpu": host_cpu(), "nproc": os.cpu_count(),
        "seconds_per_substep": med,
    }


def _adjoint_problem(ca, n, seed, steps, eps, th, dt, prec):
    """Bodies and the target snapshot: the same simulator run from the orbit
    parameters perturbed by 1%."""
    b = ca.generate(ca.IC_MERGER, n, seed)
    t = ca.Simulation(ca.generate(ca.IC_MERGER, n, seed, orbit=ca.default_orbit() * 1.01),
                      softening=eps, theta=th, dt=dt, precision=prec)
    t.advance(steps)
    target = t.state()[0]
    t.close()
    return b, target


def adjoint_runs(ca, big_n, big_steps, everies=(10, 25, 50), configs=(3, 4)):
    """Adjoint throughput.  Config 3: 8,192-body merger (seed 4), 50 substeps,
    dt 1e-3, eps 0.02, theta 0.5, FP64 walk, checkpoints every 10.  Config 4:
    big_n-body merger (seed 5), FP32 walk, theta 0.7, eps 0.003, dt 5e-4,
    big_steps forward + big_steps reverse substeps with checkpoints every K
    (segments recomputed forward) for each K, and the time-reversed leapfrog.
    Loss: mean squared distance to the target snapshot."""
    out = {}
    for cfg, name, n, seed, steps, eps, th, dt, prec, ks in [
            (3, "config3_8192_fp64_theta0.5", 8192, 4, 50, 0.02, 0.5, 1e-3, "fp64", (10,)),
            (4, f"config4_{big_n}_fp32_theta0.7", big_n, 5, big_steps, 0.003, 0.7, 5e-4, "fp32",
             everies)]:
        if cfg not in configs:
            continue
        b, target = _adjoint_problem(ca, n, seed, steps, eps, th, dt, prec)
        r = {"bodies": n, "substeps": steps}
        grads = {}
        sim = None
        for k in ks:
            # a handle of its own per K, so the device memory is that K's
            if sim is not None:
                sim.close()
            free0, total = ca.device_memory(0)
            sim = ca.Simulation(b, softening=eps, theta=th, dt=dt, precision=prec)
            t0 = time.perf_counter()
            sim.record(steps, k)
            t1 = time.perf_counter()
            xk = sim.state()[0]
            g = 2 * (xk - target) / n
            t2 = time.perf_counter()
            xb, vb, st = sim.backward(g, mode=0)
            t3 = time.perf_counter()
            free1, _ = ca.device_memory(0)
            r[f"K{k}"] = {"forward_substeps_per_s": steps / (t1 - t0),
                          "adjoint_substeps_per_s": steps / (t3 - t2),
                          "adjoint_device_bytes": float(st[1]),
                          "hbm_used_bytes": int(total - free1),
                          "hbm_added_bytes": int(free0 - free1)}
            grads[k] = xb
        t2 = time.perf_counter()
        xb, vb, st = sim.backward(g, mode=1)
        t3 = time.perf_counter()
        r["reversal"] = {"adjoint_substeps_per_s": steps / (t3 - t2),
                         "max_position_deviation": float(st[0])}
        r["reversal_vs_checkpoint_grad_rel"] = max(
            float(np.max(np.abs(grads[k] - xb)) / max(np.max(np.abs(xb)), 1e-300)) for k in ks)
        r["loss"] = float(np.sum((xk - target) ** 2) / n)
        sim.close()
        out[name] = r
    return out


def neighbor_runs(ca, bodies, radius=0.03, cap=64):
    """Batched sphere queries (the GPU octree_find_neighbors) around every body of
    the 1M merger on its FP32-path tree: kernel rate from the core's events and
    the call's rate including the PCIe copies of centres and lists."""
    sim = ca.Simulation(bodies, softening=EPS, theta=0.5, dt=DT, precision="fp32")
    pos = ca.positions(bodies)
    sim.find_neighbors(pos[:1000], radius, cap)          # geometry + buffers
    sim.set_timing(True)
    t0 = time.perf_counter()
    lists = sim.find_neighbors(pos, radius, cap)
    t1 = time.perf_counter()
    st = sim.stats()
    counts = np.fromiter((len(x) for x in lists), dtype=np.int64, count=len(lists))
    sim.close()
    return {"queries": len(pos), "radius": radius, "cap": cap,
            "queries_per_s_kernel": st.neighbor_queries / (st.ms_neighbors * 1e-3),
            "kernel_ms": st.ms_neighbors,
            "queries_per_s_with_transfers": len(pos) / (t1 - t0),
            "mean_neighbors": float(counts.mean()), "capped_fraction": float(np.mean(counts == cap))}


def sph_runs(ca, bodies, cpu=True):
    """SPH on the device tree of the 1M merger with every other star made gas:
    smoothing lengths first settle over four density passes (each moves h by
    x1.2 / x0.8 toward 20-48 neighbours), then one density and one force pass
    are timed with the core's events.  The oracle's restatement of sph.c runs
    the same two passes on the host cores from the same state, and its
    densities and accelerations are compared."""
    b = bodies.copy()
    gas = (np.arange(len(b)) % 2 == 0) & (b["type"] == 0)
    b["type"] = np.where(gas, 1, b["type"])
    b["smoothing_h"] = np.where(gas, 0.02, 0.0)
    b["internal_energy"] = np.where(gas, 1.0, 0.0)
    sim = ca.Simulation(b, softening=EPS, theta=0.5, dt=DT, precision="fp64")
    for _ in range(4):
        sim.sph_density()
    start = sim.bodies()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.sph_density()
    sim.sph_forces()
    t1 = time.perf_counter()
    st = sim.stats()
    got = sim.bodies()
    sim.close()
    ngas = int(np.sum((got["type"] == 1) & (got["mass"] > 0)))
    out = {"gas_bodies": ngas, "kernel_ms_density_forces": st.ms_sph,
           "gas_bodies_per_s_kernel": ngas / (st.ms_sph * 1e-3),
           "call_s_density_forces": t1 - t0}
    if cpu:
        from oracle.cpu_ref import CpuRef
        impl = CpuRef()
        cb = start.copy()
        pool = impl.build(cb)
        c0 = time.perf_counter()
        impl.sph_density(cb, pool)
        impl.sph_forces(cb, pool)
        c1 = time.perf_counter()
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        same = all(np.array_equal(cb[f].view(np.uint64), got[f].view(np.uint64))
                   for f in ("density", "pressure", "smoothing_h", "ax", "ay", "az"))
        out.update({"cpu_kind": impl.kind, "cpu_threads": threads,
                    "cpu_s_density_forces": c1 - c0, "cpu_gas_bodies_per_s": ngas / (c1 - c0),
                    "bit_identical_to_cpu": bool(same)})
    return out


def quasar_scene(n, seed=7):
    """Quasar-mode bodies: two spinning black holes, each with a disk of stars
    and gas (a few inside the swallow radius and the ISCO), jets with scattered
    lifetimes, some of them beyond the recycling radius.  Capacity n + n / 4."""
    rng = np.random.default_rng(seed)
    from cosmosim_amd import BODY_DTYPE
    b = np.zeros(n + n // 4, dtype=BODY_DTYPE)
    centres = np.array([[-6.0, 0.0, 0.0], [6.0, 1.0, 0.0]])
    for k in range(2):
        b["type"][k], b["mass"][k] = 2, 40.0 + 10.0 * k
        b["x"][k], b["y"][k], b["z"][k] = centres[k]
        s = rng.normal(size=3)
        b["spin_x"][k], b["spin_y"][k], b["spin_z"][k] = s / np.linalg.norm(s)
        b["accretion_rate"][k] = 30.0
    jets = 400
    m = n - 2 - jets
    host = rng.integers(0, 2, m)
    r = np.abs(rng.normal(0.0, 3.0, m)) + 0.02
    phi = rng.uniform(0, 2 * np.pi, m)
    sl = slice(2, 2 + m)
    b["x"][sl] = centres[host, 0] + r * np.cos(phi)
    b["y"][sl] = centres[host, 1] + r * np.sin(phi)
    b["z"][sl] = rng.normal(0, 0.1, m)
    vc = np.sqrt(b["mass"][host] / r) * rng.uniform(0.5, 1.1, m)
    b["vx"][sl], b["vy"][sl] = -vc * np.sin(phi), vc * np.cos(phi)
    b["mass"][sl] = 1e-4
    b["type"][sl] = np.where(rng.uniform(size=m) < 0.4, 1, 0)
    js = slice(2 + m, n)
    b["type"][js], b["mass"][js] = 3, 0.5
    b["lifetime"][js] = rng.uniform(0.0, 5.0, jets)
    jr = rng.uniform(1, 60, jets)
    b["x"][js], b["y"][js] = -6.0 + jr, rng.normal(size=jets)
    return b, n


def quasar_runs(ca, n=1_000_000, steps=10, cpu=True):
    """quasar_step on the device (accretion passes, feedback, jets) for a
    1M-body quasar scene, per step, against the oracle's restatement of
    quasar_step on one host thread for the same first steps."""
    b, n0 = quasar_scene(n)
    sim = ca.Simulation(b[:n0], softening=0.01, theta=0.5, precision="fp32")
    sim.reserve(len(b))
    cfg = ca.QuasarConfig.default(jet_count=400, jet_cap=n // 10)
    sim.quasar_step(cfg, 5e-4)                         # first-step buffers
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.quasar_step(cfg, 5e-4)
    t1 = time.perf_counter()
    out = {"bodies": n0, "steps": steps, "gpu_ms_per_step": (t1 - t0) / steps * 1e3,
           "bodies_after": sim.count(), "jet_count": cfg.jet_count}
    sim.close()
    if cpu:
        from oracle.cpu_ref import CpuRef, QuasarConfig
        c = b.copy()
        rc = QuasarConfig.default(jet_count=400, jet_cap=n // 10)
        R, kind, st = CpuRef(), "port", [1, 0]
        step = lambda k: R.quasar_step(c, k, rc, 5e-4, st)
        k = step(n0)
        c0 = time.perf_counter()
        for _ in range(3):
            k = step(k)
        c1 = time.perf_counter()
        out.update({"cpu_kind": kind, "cpu_threads": 1, "cpu_ms_per_step": (c1 - c0) / 3 * 1e3})
    return out


def _rounded(v):
    if isinstance(v, dict):
        return {k: _rounded(x) for k, x in v.items()}
    return float(f"{v:.4g}") if isinstance(v, float) else v


def host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def copy_calibration(nbytes=1 << 30, launches=5):
    """A plain device copy of known size, to scale rocprofv3's FETCH_SIZE /
    WRITE_SIZE counters: torch's elementwise negation of a float32 array
    (nbytes read and nbytes written per launch)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(launches):
        torch.neg(x, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"kernel": "neg", "bytes_read": nbytes,
                                           "bytes_written": nbytes, "launches": launches}}))


def load_traffic():
    p = os.path.join(ROOT, "profiles", "walk_fp32_list_1M_merger_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--thetas", default="0.4,0.5,0.7")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--headline-only", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-bodies", type=int, default=4_000_000)
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-everies", default="10,25,50",
                    help="checkpoint intervals K of the large adjoint configuration")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the large adjoint configuration (for profiling): prints "
                         "its figures as one JSON object")
    ap.add_argument("--no-neighbors", action="store_true")
    ap.add_argument("--no-per-step", action="store_true",
                    help="skip the per-substep subset accuracy runs")
    ap.add_argument("--per-step-steps", type=int, default=200)
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 20K-disk and 100K-merger lines")
    ap.add_argument("--no-sph", action="store_true")
    ap.add_argument("--no-quasar", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB copy used to scale the traffic counters")
    ap.add_argument("--walks", default="",
                    help="only time these walk precisions (comma list, e.g. fp64,fp64_strict) "
                         "on the 1M merger at --thetas, and print their figures as one JSON "
                         "object (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    import cosmosim_amd as ca
    everies = tuple(int(k) for k in args.adjoint_everies.split(","))
    if args.adjoint_only:
        adj = adjoint_runs(ca, args.adjoint_bodies, args.adjoint_steps, everies, configs=(4,))
        print(json.dumps({"adjoint": _rounded(adj)}))
        return
    dev_name, cus = ca.device_info(0)
    bodies = ca.generate(ca.IC_MERGER, N_BODIES, SEED)
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.choice(N_BODIES, SUBSET, replace=False)).astype(np.int32)
    thetas = [float(t) for t in args.thetas.split(",")]
    if args.walks:
        walks = {}
        for th in thetas:
            for prec in args.walks.split(","):
                progress(f"theta {th} {prec}")
                sim, walks[f"theta{th}_{prec}"] = timed_run(ca, bodies, th, prec, args.steps,
                                                            args.warmup)
                sim.close()
        print(json.dumps({"walks": _rounded(walks)}))
        return
    if 0.5 not in thetas:
        thetas.append(0.5)
    precisions = (["fp32_list"] if args.headline_only
                  else ["fp32_list", "fp32", "fp64", "fp64_strict"])

    results = {}
    head = None
    for th in thetas:
        for prec in precisions:
            progress(f"theta {th} {prec}")
            headline = (th == 0.5 and prec == "fp32_list")
            sim, r = timed_run(ca, bodies, th, prec, args.steps, args.warmup,
                               energy=(not args.no_energy and th == 0.5 and prec != "fp64"))
            if prec == "fp32_list":
                results[f"theta{th}_accuracy_end"] = accuracy(ca, sim, idx)
            sim.close()
            results[f"theta{th}_{prec}"] = r
            if headline:
                head = r
        s0 = ca.Simulation(bodies, softening=EPS, theta=th, dt=DT, precision="fp64_strict")
        results[f"theta{th}_accuracy_step0"] = accuracy(ca, s0, idx)
        s0.close()

    if not args.no_per_step:
        progress("per-step accuracy")
        results["per_step"] = {f"theta{th}_{p}": per_step_accuracy(ca, bodies, idx, th,
                                                                    args.per_step_steps, p)
                               for th in thetas for p in ("fp32_list", "fp32")}
    if not args.headline_only and not args.no_configs:
        progress("configurations 0, 1 and 4 (forward)")
        results["config0_20001_disk"] = config_line(ca, ca.IC_DISK, 20001, 1, 0.5, 0.01, 1e-3,
                                                    100, "fp64_strict")
        results["config1_100000_merger"] = config_line(ca, ca.IC_MERGER, 100000, 2, 0.5, 0.01,
                                                       1e-3, 1000, "fp64_strict")
        # config 4's forward substep (the 4M-body merger, theta 0.7, eps 0.003) with
        # each FP32 walk: the per-body one is the adjoint's, the list walk the fastest
        for prec in ("fp32_list", "fp32"):
            results[f"config4_4000000_forward_{prec}"] = config_line(
                ca, ca.IC_MERGER, 4_000_000, 5, 0.7, 0.003, 5e-4, 50, prec, energy=False)
    walk_flop = head["interactions"] * FLOP_PER_INTERACTION
    achieved = walk_flop / (head["walk_ms"] * 1e-3) / 1e12
    traffic = load_traffic()
    line = {
        "metric": METRIC,
        "value": round(head["substeps_per_s"], 3),
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "1M-body two-disk merger (seed 3), theta 0.5, eps 0.005, dt 5e-4, "
                        "FP32 walk on wave-shared interaction lists (opening test at each "
                        "wavefront's 64-body bounding box), offsets from FP64-formed "
                        "group-relative positions, FP64 state and KDK",
            "bodies": N_BODIES, "theta": 0.5, "softening": EPS, "dt": DT,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "k_walk32l",
            "bound": "valu",
            "achieved": round(achieved, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 5),
            "traffic": None if traffic is None else traffic.get("bytes_per_launch"),
            "per_launch": {"interactions": head["interactions"],
                           "flop_per_interaction": FLOP_PER_INTERACTION,
                           "avg_ms": round(head["walk_ms"], 4)},
        },
        "device": {"name": dev_name, "compute_units": cus},
    }
    # HBM fractions of the HBM-bound phases (algorithmic bytes: build reads x, y, z, m,
    # writes and reads key + index, writes 40 B per node; inside a batch the KDK is one
    # fused kick + drift per substep, reading x, v, a, m and writing x, v: 128 B per body)
    build_bytes = N_BODIES * (32 + 24) + head["nodes"] * 40
    line["phases"] = {
        "walk_ms": round(head["walk_ms"], 4), "build_ms": round(head["build_ms"], 4),
        "kdk_ms": round(head["kdk_ms"], 4),
        "build_hbm_frac": round(build_bytes / (head["build_ms"] * 1e-3) / 8e12, 4),
        "kdk_hbm_frac": round(N_BODIES * 128 / (head["kdk_ms"] * 1e-3) / 8e12, 4),
        "interactions_per_s": float(f"{head['interactions'] / (head['walk_ms'] * 1e-3):.4g}"),
    }
    free_b, total_b = ca.device_memory(0)
    line["hbm_total_bytes"] = int(total_b)
    fp64 = results.get("theta0.5_fp64")
    strict = results.get("theta0.5_fp64_strict")
    per_body = results.get("theta0.5_fp32")
    if fp64:
        line["fp32_speedup_vs_fp64"] = round(head["substeps_per_s"] / fp64["substeps_per_s"], 3)
        if per_body:
            line["fp32_per_body_speedup_vs_fp64"] = round(
                per_body["substeps_per_s"] / fp64["substeps_per_s"], 3)
    if strict:
        line["fp32_speedup_vs_fp64_strict"] = round(
            head["substeps_per_s"] / strict["substeps_per_s"], 3)
    acc = results["theta0.5_accuracy_step0"]
    line["accuracy_theta0.5"] = {k: float(f"{v:.3e}") for k, v in acc.items()}
    if not args.no_adjoint:
        progress("adjoint")
        adj = adjoint_runs(ca, args.adjoint_bodies, args.adjoint_steps, everies)
        line["adjoint"] = _rounded(adj)
        big = adj[f"config4_{args.adjoint_bodies}_fp32_theta0.7"]
        line["adjoint_substeps_per_s"] = round(big[f"K{everies[0]}"]["adjoint_substeps_per_s"], 3)
    if not args.no_neighbors:
        progress("neighbours")
        line["neighbors"] = {k: (float(f"{v:.4g}") if isinstance(v, float) else v)
                             for k, v in neighbor_runs(ca, bodies).items()}
    if not args.no_sph:
        progress("SPH")
        line["sph"] = {k: (float(f"{v:.4g}") if isinstance(v, float) else v)
                       for k, v in sph_runs(ca, bodies, cpu=not args.no_cpu).items()}
    if not args.no_quasar:
        progress("quasar")
        line["quasar"] = {k: (float(f"{v:.4g}") if isinstance(v, float) else v)
                          for k, v in quasar_runs(ca, cpu=not args.no_cpu).items()}
    if not args.no_cpu:
        progress("CPU baseline")
        cb = cpu_baseline(ca, bodies, 0.5)
        line["cpu_baseline"] = {k: (round(v, 5) if isinstance(v, float) else v)
                                for k, v in cb.items()}
        line["speedup_vs_cpu"] = round(head["substeps_per_s"] / cb["value"], 1)
    line["results"] = {k: _rounded(v) for k, v in results.items()
                       if "accuracy" not in k and k != "per_step"}
    if "per_step" in results:
        line["accuracy_per_step"] = {k: {kk: (float(f"{vv:.3e}") if isinstance(vv, float) else vv)
                                         for kk, vv in v.items()}
                                     for k, v in results["per_step"].items()}
    line["accuracy"] = {k: {kk: float(f"{vv:.3e}") for kk, vv in v.items()}
                        for k, v in results.items() if "accuracy" in k}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
