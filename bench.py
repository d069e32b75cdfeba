#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the kD-tree Barnes-Hut step on one MI355X.

Workload (the headline metric's configuration): ring distribution, N = 10M
bodies, theta = 0.3, leaf size 7, dt = 1e-3, mixed precision (FP32 far-field
and P2P with compensated FP32 sums; FP64 state, integrator and node moments).
Beside it, on the same bodies: the plain FP64 GPU step, the median / p99
relative force error of the mixed step against an FP64 direct sum on 4,096
seeded bodies, the traversal kernel's roofline from its own interaction
counters and HIP-event durations, and the reference's own C simple_sim
(Parallel/CVersion, OpenMP) timed over one whole step of the same bodies on
the host cores.  Extra legs, each in its own key: the same step at
theta = 0.5; the FP64 step at 1 M bodies over 10 steps; the 512-system
ensemble; the differentiation workload; standalone selection over 100 M
keys; the build alone on circular_orbits(10 M).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# the reference C baseline runs OpenMP in this process: idle worker threads
# sleep instead of spinning on the host cores the other legs use
os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")

from multilanguagekdtree_amd import kdtree as K  # noqa: E402

METRIC = "particle-steps/s @N=10M theta=0.3: FP32-mixed vs FP64 GPU @ median force err<=1e-6"
PEAK_FP32_VALU_TFLOPS = 157.3  # MI355X published vector FP32 peak
PEAK_FP64_VALU_TFLOPS = 78.6   # MI355X published vector FP64 peak
PEAK_HBM_GBS = 8000.0
CUS, CLOCK_HZ = 256, 2.4e9     # MI355X compute units and peak engine clock


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(bodies, n, precision, steps, warmup, theta, leaf, dt):
    with K.Simulation(n, theta=theta, dt=dt, leaf_size=leaf, precision=precision,
                      profile=True) as s:
        s.upload(bodies)
        if warmup:
            s.step(warmup)
        s.sync()
        s.reset_timings()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        t1 = time.perf_counter()
        tm = s.timings()
        # interaction counts of one more step (same tree shape, positions
        # moved by one dt): T node tests, A accepted nodes, P pairs
        s.set_params(count_interactions=1, profile=0)
        s.step(1)
        cnt = s.counters()
    wall = t1 - t0
    return dict(wall_s=wall, ms_per_step=1e3 * wall / steps,
                rate=n * steps / wall,
                build_ms=tm["build_ms"] / steps, traverse_ms=tm["traverse_ms"] / steps,
                kick_drift_ms=tm["kick_drift_ms"] / steps,
                tests=int(cnt[0]), m2p=int(cnt[1]), p2p=int(cnt[2]))


def force_error(bodies, n, theta, leaf, precision, k=4096, seed=2024):
    with K.Simulation(n, theta=theta, leaf_size=leaf, precision=precision) as s:
        s.upload(bodies)
        a = s.accel()
        idx = np.sort(np.random.default_rng(seed).choice(n, size=min(k, n), replace=False))
        ref, _ = s.direct_sum(idx)
    rel = np.linalg.norm(a[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.median(rel)), float(np.percentile(rel, 99)), float(rel.max())


def cpu_baseline(bodies, n, steps, threads, dt):
    """The reference's own C simple_sim (Parallel/CVersion/src/kdtree.c:221-260:
    build_tree on one thread, calc_accel of every body and the kick-drift as
    OpenMP parallel-for loops), run whole for `steps` steps on a copy of the
    bodies and timed around the call; without the reference build, the
    oracle's single-thread port over the same steps."""
    import oracle
    r = oracle.ref_simple_sim_timed(bodies, dt=dt, steps=steps, threads=threads)
    if r is not None:
        _, sec = r
        kind, cores = "reference", threads
        what = (f"reference C simple_sim (Parallel/CVersion kdtree.c:221-260, gcc -O2, OpenMP "
                f"{threads} threads, sequential build_tree), {steps} whole step(s) of all {n} "
                f"bodies: {sec:.2f} s")
    else:
        t0 = time.perf_counter()
        oracle.simple_sim(bodies, dt=dt, steps=steps)
        sec = time.perf_counter() - t0
        kind, cores = "port", 1
        what = f"oracle port (1 thread), {steps} whole step(s) of all {n} bodies: {sec:.2f} s"
    return dict(value=n * steps / sec, unit="particle-steps/s", cores=cores, kind=kind,
                sample=what, seconds=sec, steps=steps, n=n)


def batch_sizes(nsys, seed):
    """Per-system N, log-uniform in [2000, 200000]."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(2000), np.log(200000), nsys))).astype(np.int64)


def run_batch(nsys, steps, loop_steps, theta, leaf, dt, seed):
    """The ensemble: nsys ring systems (system i from seed seed + 1 + i) built
    and stepped together by one graph-replayed batch step, against a loop
    that steps each system in its own context (kd_step per system)."""
    sizes = batch_sizes(nsys, seed)
    total = int(sizes.sum())
    bodies = K.ensemble(sizes, seed0=seed + 1)
    with K.Batch(sizes, theta=theta, dt=dt, leaf_size=leaf, precision=K.KD_PREC_MIXED) as b:
        b.upload(bodies)
        tune = b.autotune(reps=2)
        log(f"[bench] batch: autotuned walk {tune}")
        b.step(1)  # captures the two step graphs
        t0 = time.perf_counter()
        b.step(steps)
        t1 = time.perf_counter()
    rate = total * steps / (t1 - t0)
    sims = []
    try:
        for i, nn in enumerate(sizes):
            s = K.Simulation(int(nn), theta=theta, dt=dt, leaf_size=leaf,
                             precision=K.KD_PREC_MIXED)
            s.upload({k: v[b.offsets[i]:b.offsets[i + 1]] for k, v in bodies.items()})
            sims.append(s)
        for s in sims:
            s.step(1)  # warm-up
        t2 = time.perf_counter()
        for s in sims:
            s.step(loop_steps)
        t3 = time.perf_counter()
    finally:
        for s in sims:
            s.close()
    loop_rate = total * loop_steps / (t3 - t2)
    return dict(systems=nsys, bodies=total, n_min=int(sizes.min()), n_max=int(sizes.max()),
                steps=steps, value=rate, ms_per_step=1e3 * (t1 - t0) / steps,
                per_system_loop=dict(value=loop_rate, steps=loop_steps,
                                     ms_per_step=1e3 * (t3 - t2) / loop_steps),
                speedup_vs_loop=rate / loop_rate, walk_shape=tune,
                unit="particle-steps/s", precision="mixed", graph=True)


def run_adjoint(n, steps, every, theta, leaf, dt, seed, ndir, cpu_n):
    """Differentiation workload: gradients of sum |p_N - p_ref|^2 (p_ref: the
    FP64 run from velocities + N(0, 1e-4)) with respect to initial velocities
    and masses, checkpoint every `every` steps, in both precisions; each
    directional derivative is compared with central differences (eps 1e-7)
    of the FP64 forward along `ndir` seeded random directions (the mixed
    forward reads FP32 positions, so a 1e-7 perturbation is below its
    resolution and its own differences are no derivative), and (at cpu_n
    bodies) the gradients with the oracle's FP64 reference gradient."""
    import torch
    from multilanguagekdtree_amd.autograd import Adjoint
    import oracle

    def cuda(a):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")

    P = lambda s: np.stack([s["x"], s["y"], s["z"]], 1)  # noqa: E731
    V = lambda s: np.stack([s["vx"], s["vy"], s["vz"]], 1)  # noqa: E731

    def final_positions(b, prec):
        with K.Simulation(len(b["x"]), theta=theta, dt=dt, leaf_size=leaf, precision=prec) as s:
            s.upload(b)
            s.step(steps)
            return P(s.download())

    def grads(b, pref, prec):
        sim = Adjoint(len(b["x"]), steps, every, theta=theta, dt=dt, leaf_size=leaf,
                      precision=prec)
        x0, v0, m = cuda(P(b)), cuda(V(b)), cuda(b["m"])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        pN, _ = sim.forward(x0, v0, m)
        t1 = time.perf_counter()
        gx, gv, gm = sim.backward(2 * (pN - cuda(pref)), None)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        sim.close()
        return gv.cpu().numpy(), gm.cpu().numpy(), t1 - t0, t2 - t1, pN.cpu().numpy()

    rng = np.random.default_rng(seed + 7)
    b = K.ring(n, seed=seed + 3)
    ref = dict(b)
    for k in ("vx", "vy", "vz"):
        ref[k] = b[k] + rng.normal(size=n) * 1e-4
    pref = final_positions(ref, K.KD_PREC_FP64)
    eps, fds = 1e-7, []
    t0 = time.perf_counter()
    for _ in range(ndir):
        dv = rng.normal(size=(n, 3))
        dm = b["m"] * rng.normal(size=n)

        def loss(sgn):
            c = dict(b)
            c["vx"], c["vy"], c["vz"] = (V(b) + sgn * eps * dv).T
            c["m"] = b["m"] + sgn * eps * dm
            return ((final_positions(c, K.KD_PREC_FP64) - pref) ** 2).sum()
        fds.append((dv, dm, (loss(1) - loss(-1)) / (2 * eps)))
    log(f"[bench] adjoint: {ndir} FP64 central differences in {time.perf_counter() - t0:.1f} s")
    p64 = final_positions(b, K.KD_PREC_FP64)
    out = {"n": n, "steps": steps, "checkpoint_every": every, "fd_directions": ndir,
           "fd_eps": eps, "fd_forward": "fp64"}
    for name, prec in (("mixed", K.KD_PREC_MIXED), ("fp64", K.KD_PREC_FP64)):
        grads(b, pref, prec)  # warm-up
        gv, gm, tf, tb, pN = grads(b, pref, prec)
        devs = [abs(fd - float((gv * dv).sum() + (gm * dm).sum())) /
                abs(float((gv * dv).sum() + (gm * dm).sum())) for dv, dm, fd in fds]
        out[name] = {"forward_s": tf, "backward_s": tb,
                     "forward_particle_steps_per_s": n * steps / tf,
                     "fwd_bwd_particle_steps_per_s": n * steps / (tf + tb),
                     "fd_rel_dev_median": float(np.median(devs)),
                     "fd_rel_dev_max": float(np.max(devs))}
        if name == "mixed":
            # first-order size of the deviation: the two forwards' final
            # positions differ by d, and the gradient is 2 J^T (p_N - p_ref)
            out[name]["traj_shift_over_residual"] = float(
                np.linalg.norm(pN - p64) / np.linalg.norm(p64 - pref))
        log(f"[bench] adjoint {name}: {out[name]}")
    # FP64 CPU reference gradient (the oracle, frozen topology) at cpu_n bodies
    bc = K.ring(cpu_n, seed=seed + 5)
    refc = dict(bc)
    for k in ("vx", "vy", "vz"):
        refc[k] = bc[k] + rng.normal(size=cpu_n) * 1e-4
    prc, _ = oracle.adj_forward(refc, dt, steps, theta, leaf)
    fin, tape = oracle.adj_forward(bc, dt, steps, theta, leaf)
    _, ov, om = oracle.adj_backward(tape, 2 * (P(fin) - P(prc)), np.zeros((cpu_n, 3)))
    del tape
    cmp = {}
    for name, prec in (("mixed", K.KD_PREC_MIXED), ("fp64", K.KD_PREC_FP64)):
        gv, gm, _, _, _ = grads(bc, P(prc), prec)
        cmp[name] = {"v_rel": float(np.linalg.norm(gv - ov) / np.linalg.norm(ov)),
                     "m_rel": float(np.linalg.norm(gm - om) / np.linalg.norm(om))}
    out["vs_cpu_fp64_gradient"] = dict(n=cpu_n, steps=steps, **cmp)
    log(f"[bench] adjoint vs CPU FP64 gradient: {out['vs_cpu_fp64_gradient']}")
    return out


def run_build_only(n, leaf, seed, reps):
    """Build-only benchmark after Parallel/RustVersion/src/bin/bench_build_tree.rs:6-41:
    circular_orbits(n) (n + 1 bodies), tree build alone; GPU build from the
    initial-condition order (first) and from the previous tree order (steady
    state), beside the reference C sequential build_tree of the same bodies."""
    import ctypes as C
    import oracle
    b = K.circular_orbits(n, seed=seed)
    nb = len(b["x"])
    with K.Simulation(nb, leaf_size=leaf) as s:
        s.upload(b)
        t0 = time.perf_counter()
        s.build()
        t1 = time.perf_counter()
        times = []
        for _ in range(reps):
            t2 = time.perf_counter()
            s.build()
            times.append(time.perf_counter() - t2)
    out = {"bodies": nb, "leaf_size": leaf, "gpu_first_build_ms": 1e3 * (t1 - t0),
           "gpu_build_ms_median": 1e3 * float(np.median(times)), "reps": reps}
    R = oracle.ref_lib("par")
    if R is not None:
        f = {k: np.ascontiguousarray(b[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "m")}
        d = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
        idx = np.zeros(1, np.int64)
        acc = np.empty(3)
        t = np.zeros(3)
        R.ref_cpu_sample(nb, d(f["x"]), d(f["y"]), d(f["z"]), d(f["vx"]), d(f["vy"]), d(f["vz"]),
                         d(f["m"]), 1, idx.ctypes.data_as(C.POINTER(C.c_int64)), 1, 1e-3,
                         d(acc), d(t))
        out["cpu_reference_build_s"] = float(t[0])
        out["cpu_reference"] = "Parallel/CVersion build_tree (sequential, leaf 7, gcc -O2), 1 thread"
        out["speedup_vs_cpu"] = float(t[0]) / (out["gpu_build_ms_median"] * 1e-3)
    return out


def run_select(n, seed, reps):
    """Standalone selection after Parallel/RustVersion/src/bin/bench_quickstat.rs:6-49:
    n uniform values, indices 0..n, one random goal; GPU selection on
    device-resident arrays beside the oracle's sequential quickselect."""
    import ctypes as C
    import torch
    import oracle
    from multilanguagekdtree_amd import _lib
    rng = np.random.default_rng(seed)
    vals = rng.random(n)
    goal = int(rng.integers(0, n))
    L = _lib.lib()
    dv = torch.as_tensor(vals, device="cuda")
    base = torch.arange(n, dtype=torch.int64, device="cuda")
    di = base.clone()
    wsb = int(L.kd_quickstat_workspace_bytes(n))
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    times = []
    for r in range(reps + 1):
        di.copy_(base)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.kd_quickstat_index_device(n, C.c_void_p(di.data_ptr()), goal,
                                               C.c_void_p(dv.data_ptr()), n,
                                               C.c_void_p(ws.data_ptr()), wsb))
        t1 = time.perf_counter()
        if r:
            times.append(t1 - t0)
    pivot = int(di[goal].item())
    # rank check: exactly `goal` values below the pivot's (values are distinct)
    assert int((dv < dv[pivot]).sum().item()) == goal
    del dv, di, ws, base
    t0 = time.perf_counter()
    ref = oracle.quickstat_index(np.arange(n), goal, vals)
    t_cpu = time.perf_counter() - t0
    assert int(ref[goal]) == pivot
    return {"n": n, "goal": goal, "gpu_s_median": float(np.median(times)), "reps": reps,
            "cpu_oracle_s": t_cpu, "cpu": "oracle quickselect (C, 1 thread, seeded pivots)",
            "speedup_vs_cpu": t_cpu / float(np.median(times))}


def run_child(argv):
    """Run one leg of this script in a child process; its JSON line, or an
    error record if the child failed."""
    import subprocess
    r = subprocess.run([sys.executable, str(Path(__file__).resolve())] + argv,
                       capture_output=True, text=True)
    sys.stderr.write(r.stderr)
    if r.returncode != 0:
        return {"error": f"child exited with {r.returncode}", "argv": argv}
    return json.loads(r.stdout.strip().splitlines()[-1])


def tree_depth(n, leaf):
    """Levels below the root of the tree of n bodies (every split at floor(len/2))."""
    d, mx = 0, n
    while mx > leaf:
        mx = (mx + 1) // 2
        d += 1
    return d


def compulsory_bytes(n, leaf):
    """Compulsory HBM bytes per particle-step: read p, v, m and write p, v in
    FP64 (104 B), 16 B per build level (8 B key + 4 B index in + 4 B out) and
    ~20 B of node records."""
    return 104 + 16 * tree_depth(n, leaf) + 20


def step_leg(bodies, n, steps, warmup, theta, leaf, dt, precs=("fp64", "mixed"), err=True):
    """One workload timed in the given precisions (+ the mixed force error)."""
    out = {"n": n, "theta": theta, "leaf_size": leaf, "steps": steps}
    for name in precs:
        prec = K.KD_PREC_MIXED if name == "mixed" else K.KD_PREC_FP64
        r = run_gpu(bodies, n, prec, steps, warmup, theta, leaf, dt)
        out[name] = {"value": r["rate"], "ms_per_step": r["ms_per_step"],
                     "phases_ms": {k: round(r[k], 3)
                                   for k in ("build_ms", "traverse_ms", "kick_drift_ms")},
                     "algorithmic_hbm_gbps": round(compulsory_bytes(n, leaf) * r["rate"] / 1e9, 1)}
        log(f"[bench] {name} n={n} theta={theta}: {json.dumps(out[name])}")
    if "fp64" in out and "mixed" in out:
        out["speedup_mixed_vs_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    if err and "mixed" in out:
        med, p99, mx = force_error(bodies, n, theta, leaf, K.KD_PREC_MIXED)
        out["mixed"]["force_error"] = dict(median=med, p99=p99, max=mx, samples=4096)
        log(f"[bench] force error n={n} theta={theta}: {out['mixed']['force_error']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # two warm-up steps: the first builds from the upload's order (all three
    # lists, kernel-by-kernel levels), the second is the first steady build
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--theta", type=float, default=0.3)
    ap.add_argument("--leaf", type=int, default=7)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    # whole steps of the reference C simple_sim on the headline bodies (one
    # 10 M step takes about a minute on 16 host cores) and the 10 steps of the
    # 1 M FP64 workload (about 25 s)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--cpu-config1-steps", type=int, default=10)
    ap.add_argument("--cpu-repeats", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--only", choices=["mixed", "fp64", "both", "batch", "adjoint", "copy",
                                       "build", "select", "cpu"], default="both")
    ap.add_argument("--select-n", type=int, default=100_000_000)
    ap.add_argument("--copy-mib", type=int, default=1024)
    # MAX_PARTS of the crate the build-only benchmark comes from
    # (Parallel/RustVersion/src/array_kd_tree.rs:14)
    ap.add_argument("--build-leaf", type=int, default=8)
    ap.add_argument("--adj-n", type=int, default=100_000)
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-every", type=int, default=20)
    ap.add_argument("--adj-dirs", type=int, default=64)
    ap.add_argument("--adj-cpu-n", type=int, default=2000)
    ap.add_argument("--batch-systems", type=int, default=512)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-loop-steps", type=int, default=5)
    ap.add_argument("--config1-n", type=int, default=1_000_000)
    ap.add_argument("--config1-steps", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline legs only: no theta 0.5, 1 M, adjoint, selection or "
                         "build-only legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    n, dt = args.n, 1e-3
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", 0)) or (os.cpu_count() or 1)
    if args.only == "select":
        print(json.dumps(run_select(args.select_n, args.seed, 3)), flush=True)
        return
    if args.only == "build":
        rb = run_build_only(args.n, args.build_leaf, args.seed, 5)
        print(json.dumps(rb), flush=True)
        return
    if args.only == "copy":
        # a plain device-to-device copy of known size: the calibration run for
        # the FETCH_SIZE / WRITE_SIZE counters (one kernel per copy)
        import torch
        nbytes = args.copy_mib << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "copies": 3}), flush=True)
        return
    if args.only == "adjoint":
        log(f"[bench] adjoint: n={args.adj_n} steps={args.adj_steps}")
        ra = run_adjoint(args.adj_n, args.adj_steps, args.adj_every, args.theta, args.leaf, dt,
                         args.seed, args.adj_dirs, args.adj_cpu_n)
        print(json.dumps(ra), flush=True)
        return
    if args.only == "batch":
        log(f"[bench] batch of {args.batch_systems} systems, {args.batch_steps} steps")
        rb = run_batch(args.batch_systems, args.batch_steps, args.batch_loop_steps, args.theta,
                       args.leaf, dt, args.seed)
        print(json.dumps(rb), flush=True)
        return
    if args.only == "cpu":
        # the reference C simple_sim alone: median of --cpu-repeats runs
        out = {}
        for nn, st in ((n, args.cpu_steps), (args.config1_n, args.config1_steps)):
            if st <= 0:
                continue
            b = K.ring(nn, seed=args.seed)
            runs = []
            for rep in range(args.cpu_repeats):
                runs.append(cpu_baseline(b, nn, st, threads, dt))
                log(f"[bench] cpu n={nn} steps={st} run {rep + 1}: {runs[-1]['seconds']:.2f} s")
            vals = sorted(r["value"] for r in runs)
            out[f"n{nn}_steps{st}"] = dict(runs[0], value=vals[len(vals) // 2],
                                          values=vals, repeats=args.cpu_repeats)
            log(f"[bench] cpu n={nn} steps={st}: {out[f'n{nn}_steps{st}']}")
        print(json.dumps(out), flush=True)
        return
    log(f"[bench] generating ring ICs n={n} seed={args.seed}")
    bodies = K.ring(n, seed=args.seed)
    res = {}
    for name, prec in (("fp64", K.KD_PREC_FP64), ("mixed", K.KD_PREC_MIXED)):
        if args.only not in ("both", name):
            continue
        log(f"[bench] {name}: warmup {args.warmup}, timing {args.steps} steps")
        res[name] = run_gpu(bodies, n, prec, args.steps, args.warmup, args.theta, args.leaf, dt)
        log(f"[bench] {name}: {json.dumps(res[name])}")
    main_name = "mixed" if "mixed" in res else "fp64"
    r = res[main_name]

    err = None
    if not args.no_error:
        log("[bench] force error vs FP64 direct sum on 4096 bodies")
        prec = K.KD_PREC_MIXED if main_name == "mixed" else K.KD_PREC_FP64
        med, p99, mx = force_error(bodies, n, args.theta, args.leaf, prec)
        err = dict(median=med, p99=p99, max=mx, samples=4096)
        log(f"[bench] error {err}")

    # roofline of the dominant kernel (the traversal, which also applies the
    # kick-drift to its bodies): algorithmic FLOPs per launch = 10 T + 9 A +
    # 18 P over all bodies, from the kernel's counters, + 12 per body for the
    # kick-drift (v += dt a; p += dt v)
    flops = 10 * r["tests"] + 9 * r["m2p"] + 18 * r["p2p"] + 12 * n
    achieved = flops / (r["traverse_ms"] * 1e-3) / 1e12
    peak = PEAK_FP32_VALU_TFLOPS if main_name == "mixed" else PEAK_FP64_VALU_TFLOPS
    # HBM bytes per launch of the walk: FETCH_SIZE and WRITE_SIZE counter
    # passes, each scaled by what the counter reads for a plain copy of known
    # size (profiles/traffic_ring10M_mixed.md)
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        try:
            t = json.loads(tf.read_text()).get(f"traverse_{main_name}_n{n}")
            traffic = None if t is None else float(t["bytes"])
        except Exception:
            traffic = None
    # vector-ALU counters of the walk from the committed counter pass
    # (profiles/valu_ring10M_mixed.md): VALU instructions per wavefront and
    # VALUUtilization (the share of lanes doing work in issued instructions);
    # the issue rate is per CU and clock over this run's walk duration
    valu = None
    vf = ROOT / "profiles" / "valu.json"
    if vf.exists() and main_name == "mixed":
        try:
            valu = json.loads(vf.read_text()).get(f"traverse_{main_name}_n{n}")
        except Exception:
            valu = None
    issue = None
    if valu is not None:
        waves = (n + 63) // 64
        issue = valu["valu_insts_per_wave"] * waves / (r["traverse_ms"] * 1e-3 * CUS * CLOCK_HZ)
    roofline = dict(bound="valu", achieved=round(achieved, 3), peak=peak, unit="TFLOP/s",
                    frac=round(achieved / peak, 4), traffic=traffic,
                    # the walk's kernel as rocprofv3 names it (profiles/step_kernels_*.md)
                    kernel=("kd::k_walk<kd::F32, false, false, true>" if main_name == "mixed"
                            else "kd::k_walk_f64<false, true>"),
                    kernel_ms=round(r["traverse_ms"], 3),
                    hbm_gbps=(None if traffic is None
                              else round(traffic / (r["traverse_ms"] * 1e-3) / 1e9, 1)),
                    valu_insts_per_wave=None if valu is None else round(valu["valu_insts_per_wave"]),
                    valu_issue_per_cu_cycle=None if issue is None else round(issue, 3),
                    valu_utilization_pct=(None if valu is None
                                          else round(valu["valu_utilization_pct"], 1)),
                    flops_per_launch=flops,
                    per_particle=dict(tests=r["tests"] / n, m2p=r["m2p"] / n, p2p=r["p2p"] / n))

    extra = {}
    if not args.no_extra:
        log("[bench] theta 0.5 leg")
        extra["theta_0.5"] = step_leg(bodies, n, args.steps, args.warmup, 0.5, args.leaf, dt)
    del bodies

    # the CPU legs run in a child process of their own: inside this process,
    # after the GPU legs, the same reference code measured 1.7 to 5 times
    # slower than in a fresh process
    cpu = c1 = None
    if not args.no_cpu_baseline:
        st1 = 0 if args.no_extra else args.cpu_config1_steps
        log(f"[bench] CPU baseline: reference C simple_sim, {threads} threads, "
            f"{args.cpu_steps} whole step(s) at n={n}"
            + (f", {st1} at n={args.config1_n}" if st1 else ""))
        rc = run_child(["--only", "cpu", "--n", str(n), "--cpu-steps", str(args.cpu_steps),
                        "--config1-n", str(args.config1_n), "--config1-steps", str(st1),
                        "--cpu-threads", str(threads), "--seed", str(args.seed)])
        cpu = rc.get(f"n{n}_steps{args.cpu_steps}")
        c1 = rc.get(f"n{args.config1_n}_steps{st1}")
        log(f"[bench] cpu {cpu}")

    if not args.no_extra:
        n1 = args.config1_n
        log(f"[bench] FP64 leg n={n1}, {args.config1_steps} steps")
        b1 = K.ring(n1, seed=args.seed)
        extra["config1_fp64"] = step_leg(b1, n1, args.config1_steps, args.warmup, args.theta,
                                         args.leaf, dt, precs=("fp64",), err=False)
        if c1 is not None:
            extra["config1_fp64"]["cpu_reference"] = c1
            extra["config1_fp64"]["speedup_vs_cpu"] = (
                extra["config1_fp64"]["fp64"]["value"] / c1["value"])
            log(f"[bench] cpu n={n1}: {c1}")
        del b1

    batch = None
    if not args.no_batch:
        log(f"[bench] batch of {args.batch_systems} systems, {args.batch_steps} steps")
        batch = run_batch(args.batch_systems, args.batch_steps, args.batch_loop_steps,
                          args.theta, args.leaf, dt, args.seed)
        log(f"[bench] batch {json.dumps(batch)}")

    if not args.no_extra:
        # the legs that use torch run in a child process of their own (the
        # parent has created and destroyed hundreds of library contexts by now)
        log(f"[bench] adjoint: n={args.adj_n} steps={args.adj_steps}")
        extra["adjoint"] = run_child(["--only", "adjoint", "--adj-n", str(args.adj_n),
                                      "--adj-steps", str(args.adj_steps),
                                      "--adj-every", str(args.adj_every),
                                      "--adj-dirs", str(args.adj_dirs),
                                      "--adj-cpu-n", str(args.adj_cpu_n),
                                      "--theta", str(args.theta), "--leaf", str(args.leaf),
                                      "--seed", str(args.seed)])
        log("[bench] selection leg")
        extra["select"] = run_child(["--only", "select", "--select-n", str(args.select_n),
                                     "--seed", str(args.seed)])
        log(f"[bench] select {extra['select']}")
        log("[bench] build-only leg")
        extra["build_only"] = run_child(["--only", "build", "--n", str(n),
                                         "--build-leaf", str(args.build_leaf),
                                         "--seed", str(args.seed)])
        log(f"[bench] build only {extra['build_only']}")

    out = {
        "metric": (METRIC if (n, args.theta) == (10_000_000, 0.3) else
                   f"particle-steps/s @N={n} theta={args.theta}: FP32-mixed vs FP64 GPU "
                   "@ median force err<=1e-6"),
        "value": r["rate"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_name == "mixed" else "f64",
        "data": "synthetic (seeded ring ICs generated in-process)",
        "config": {"workload": f"kd-tree Barnes-Hut step, ring N={n}, theta={args.theta}, "
                               f"leaf {args.leaf}, dt {dt}, {main_name} precision",
                   "n": n, "theta": args.theta, "leaf_size": args.leaf, "dt": dt,
                   "precision": main_name, "seed": args.seed},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "phases_ms": {k: round(r[k], 3) for k in ("build_ms", "traverse_ms", "kick_drift_ms")},
        "algorithmic_hbm_gbps": round(compulsory_bytes(n, args.leaf) * r["rate"] / 1e9, 1),
        "force_error": err,
        "batch": batch,
    }
    if "fp64" in res and "mixed" in res:
        out["fp64_gpu"] = {"value": res["fp64"]["rate"], "ms_per_step": res["fp64"]["ms_per_step"],
                           "phases_ms": {k: round(res["fp64"][k], 3)
                                         for k in ("build_ms", "traverse_ms", "kick_drift_ms")}}
        out["speedup_mixed_vs_fp64"] = res["mixed"]["rate"] / res["fp64"]["rate"]
    if cpu is not None:
        out["speedup_vs_cpu"] = r["rate"] / cpu["value"]
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
