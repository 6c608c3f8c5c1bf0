#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-frame N-body step on one MI355X.

Workload (N=1): the plain GPU configuration -- one collision of 2 x 131,072
particles (70/10/20 disc/bulge/halo), FP32, theta = 0.5, softening 0.02,
dt = 0.005, leaf 4, groups of at most 96 each walked and evaluated by a block of
four waves, quadrupole group traversal.  Inputs are generated on the
host and resident in HBM before the timed region; each timed step is a full
step (bounds, Morton, sort, permute, tree build, aggregation, traversal,
kick-drift), replayed from a hipGraph.

Prints one JSON line.  `roofline` is the traversal kernel: 20 flop per
interaction (interactions counted exactly by the kernel) over its average
duration from HIP events on the engine stream, against the MI355X FP32 peak
(157.3 TFLOP/s, the matrix-core figure; `bound` is "mfma", the compute class,
although the kernel issues vector FP32).
`cpu_baseline` times the reference's own CPU octree step (compiled from its
sources into oracle/_ref) on a bounded sample of the same input on the host
cores; the same leg checks the GPU accelerations against an FP64 direct sum.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "galaxy-collision-simulator_amd"))

PEAK_FP32_TFLOPS = 157.3
FLOP_PER_INTERACTION = 20
# FP32 operations the kernel executes per interaction (a multiply-add counted as
# two, the reciprocal square root as one): softened quadrupole node 52,
# particle 19 (interact_n / interact_p in csrc/gcs_traverse.hpp).
EXEC_FLOP_NODE = 52
EXEC_FLOP_PARTICLE = 19
PEAK_FP32_VECTOR_TFLOPS = 78.6  # vector FP32 without packed operations
PEAK_HBM_GBS = 8000.0
TREE_BYTES_PER_PARTICLE = 376  # Morton 24 + sort 24 + permute 68 + init 88 + build 24 + agg 148

# HBM bytes per traversal launch, from rocprofv3 FETCH_SIZE (x2.000) and
# WRITE_SIZE (x1.000) passes scaled by a 256 MiB copy (profiles/
# traversal_counters_plain-2x131072.md); valid for the configuration it was
# measured on: (workload, leaf, block, waves per SIMD, group, team).
TRAVERSE_TRAFFIC = {("plain-2x131072", 4, 256, 4, 96, 4): 137057843}

CONFIGS = {
    # name: (particles per galaxy, eps, dt, theta)
    "plain-2x131072": (131072, 0.02, 0.005, 0.5),
    "plain-2x8192": (8192, 0.05, 0.01, 0.5),
    "precision-2x1048576": (1048576, 0.02, 0.005, 0.5),
}


def accuracy_check(stars, p, acc_gpu, targets):
    """CPU checker of the GPU accelerations: FP64 direct sum (the oracle) on
    `targets` evenly spaced particles; relative RMS and p99 per-particle
    relative error."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    n = len(stars)
    tgt = np.linspace(0, n - 1, targets).astype(np.int64)
    ref = oracle.direct_sum_f64(stars["x"], stars["y"], stars["z"], stars["mass"], p.G,
                                float(p.eps2), tgt)
    d = acc_gpu[tgt, :3].astype(np.float64) - ref[:, :3]
    err = float(np.sqrt((d ** 2).sum() / (ref[:, :3] ** 2).sum()))
    rel = np.linalg.norm(d, axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    return {"rel_rms_vs_fp64": err, "p99_rel_vs_fp64": float(np.quantile(rel, 0.99)),
            "targets": len(tgt)}


def cpu_octree_rate(stars, p, sample_seconds, max_steps=50, fast=False):
    """The reference's compiled CPU octree frame (oracle/_ref; fast: the build
    with -ffast-math, as the reference's README builds it) timed on the host
    cores; particle-steps/s."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    kind = "reference" if oracle.ref_available(fast) else "port"
    step = (lambda s: oracle.ref_octree_step(s, p.G, float(p.eps2), p.dt, fast=fast)) \
        if kind == "reference" \
        else (lambda s: oracle.octree_step(s, p.G, float(p.eps2), p.dt, 0.8, 32))
    s = stars
    s, _, _ = step(s)  # warm-up
    t0 = time.perf_counter()
    k = 0
    while True:
        s, _, _ = step(s)
        k += 1
        if time.perf_counter() - t0 > sample_seconds or k >= max_steps:
            break
    return len(stars) * k / (time.perf_counter() - t0), k, kind


def cpu_baseline(stars, p, acc_gpu, sample_seconds):
    """Times the reference CPU octree on the host and checks the GPU
    accelerations against an FP64 direct sum on a subsample."""
    n = len(stars)
    check = accuracy_check(stars, p, acc_gpu, 4096)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rate, k, kind = cpu_octree_rate(stars, p, sample_seconds)
    out = {
        "value": rate,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": kind,
        "sample": "%d steps of the reference CPU octree frame (theta=0.8, leaf 32, as compiled "
                  "in gravity_aggregator.cpp/engine.cpp) on the same %d-particle input, "
                  "OMP_NUM_THREADS=%d" % (k, n, threads),
    }
    import oracle
    if kind == "reference" and oracle.ref_available(True):
        rf, kf, _ = cpu_octree_rate(stars, p, sample_seconds / 2, fast=True)
        out["fast_math"] = {"value": rf, "steps": kf,
                            "build": "g++ -O3 -fopenmp -ffast-math -march=x86-64-v3"}
    return out, check


def cpu_equal_accuracy(stars, p, gpu_value, thetas=(0.6, 0.5, 0.4, 0.35, 0.3, 0.25)):
    """The CPU octree (the oracle's restatement of the reference octree, with
    the opening angle as a parameter, leaf cap 32, host threads) at the first
    angle whose accelerations meet 1e-3 relative RMS against the same FP64
    check as the GPU: that angle, its error, one frame's time and the GPU's
    speed-up over it.  The reference hard-codes theta = 0.8 (the cpu_baseline
    above), which misses 1e-3 by an order of magnitude."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    n = len(stars)
    tried = []
    for th in (0.8,) + tuple(thetas):
        t0 = time.perf_counter()
        s1, acc, _ = oracle.octree_step(stars, p.G, float(p.eps2), p.dt, th, 32,
                                        integrate=False, want_acc=True)
        el = time.perf_counter() - t0
        chk = accuracy_check(s1, p, acc, 4096)
        tried.append({"theta": th, "rel_rms_vs_fp64": chk["rel_rms_vs_fp64"],
                      "frame_s": el, "particle_steps_per_s": n / el})
        if th < 0.8 and chk["rel_rms_vs_fp64"] <= 1e-3:
            return {"theta": th, "leaf": 32, "kind": "port",
                    "rel_rms_vs_fp64": chk["rel_rms_vs_fp64"],
                    "value": n / el, "unit": "particle-steps/s",
                    "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                    "speedup_vs_cpu": gpu_value / (n / el), "tried": tried}
    return {"theta": None, "tried": tried}


def precision_study(steps, targets, cpu_seconds, seed=1003):
    """The precision form: one collision of 2 x 1,048,576 particles, eps 0.02,
    dt 0.005.  For theta 0.5 and 0.7: the initial accelerations against an
    FP64 direct sum on `targets` particles (CPU checker), and the relative
    total-energy drift over `steps` steps with plain FP32 and with compensated
    positions (FP64 diagnostics on the GPU, first and last step).  Also the
    GPU step rate next to the reference CPU octree on the same input."""
    import numpy as np
    import gcs_amd

    stars = gcs_amd.collision(1048576, seed)
    n = len(stars)
    eps2 = float(np.float32(0.02) ** 2)
    out = {"particles": n, "steps": steps, "runs": []}
    for theta in (0.5, 0.7):
        base = gcs_amd.params(G=1.0, eps2=eps2, dt=0.005, theta=theta, use_graph=1)
        with gcs_amd.Engine(n) as e:
            e.load(stars)
            base, _ = e.autotune(base, reps=2)
            e.compute_accels(base)
            acc = e.accels()
            s0 = e.stars()
        run = {"theta": theta, "leaf": base.leaf_size, "block": base.block_size,
               "waves_per_simd": base.waves_per_simd, "group": base.group_size,
               "team": base.team}
        run.update(accuracy_check(s0, base, acc, targets))
        for prec, name in ((gcs_amd.PREC_PLAIN, "plain"),
                           (gcs_amd.PREC_COMPENSATED, "compensated")):
            kw = dict(G=1.0, eps2=eps2, dt=0.005, theta=theta, leaf_size=base.leaf_size,
                      block_size=base.block_size, waves_per_simd=base.waves_per_simd,
                      group_size=base.group_size, team=base.team, precision=prec,
                      use_graph=1)
            pd = gcs_amd.params(diagnostics=1, **kw)
            pn = gcs_amd.params(diagnostics=0, **kw)
            with gcs_amd.Engine(n) as e:
                e.load(stars)
                e.step(pd, 1)
                d0 = e.diagnostics()
                t0 = time.perf_counter()
                e.step(pn, steps - 2, sync=False)
                e.sync()
                el = time.perf_counter() - t0
                e.step(pd, 1)
                d1 = e.diagnostics()
            e0, e1 = d0[0] + d0[1], d1[0] + d1[1]
            run[name] = {"energy_first": e0, "energy_last": e1,
                         "energy_drift_rel": (e1 - e0) / abs(e0),
                         "momentum_change": float(np.linalg.norm(d1[2:5] - d0[2:5])),
                         "ms_per_step": 1e3 * el / (steps - 2)}
        out["runs"].append(run)
        print("precision theta %.1f: %s" % (theta, json.dumps(run)), file=sys.stderr, flush=True)
    plain05 = out["runs"][0]["plain"]["ms_per_step"]
    cpu_p = gcs_amd.params(G=1.0, eps2=eps2, dt=0.005)
    rate, k, kind = cpu_octree_rate(stars, cpu_p, cpu_seconds, max_steps=5)
    out["gpu_particle_steps_per_s"] = n / (plain05 * 1e-3)
    out["cpu_octree"] = {"value": rate, "unit": "particle-steps/s", "kind": kind, "steps": k,
                         "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    out["speedup_vs_cpu"] = out["gpu_particle_steps_per_s"] / rate
    return out


def batched_run(n_inst, steps, warmup, leaf, seed=1004, plain_steps=10, plain_sample=32):
    """Batched form: n_inst collisions (N log-uniform in [4096, 65536]) advanced
    together as one forest, against the same collisions run one at a time
    through the plain path (a sample of plain_sample of them, plain_steps steps
    each, extrapolated per particle-step).  Each form runs at the settings the
    autotuner picks for it: the forest's for the batched form, a single
    collision of the median size's for the plain path.  Returns a dict."""
    import numpy as np
    import gcs_amd

    specs = [gcs_amd.random_spec(seed, k, 4096, 65536) for k in range(n_inst)]
    counts = [sp.n_a + sp.n_b for sp in specs]
    total = int(sum(counts))
    p = gcs_amd.params(G=1.0, eps2=float(np.float32(0.02) ** 2), dt=0.005, theta=0.5,
                       leaf_size=leaf, precision=gcs_amd.PREC_COMPENSATED, use_graph=1)
    e = gcs_amd.Engine(total, n_inst)
    e.load_generated(specs)  # initial conditions generated on the device
    p, _ = e.autotune(p, reps=2)
    e.step(p, warmup)
    t0 = time.perf_counter()
    e.step(p, steps, sync=False)
    e.sync()
    el = time.perf_counter() - t0
    e.close()
    batched = total * steps / el
    # One at a time: each collision alone (fresh engine, graph-replayed steps),
    # at the settings tuned on the median-size collision.
    med = sorted(zip(counts, range(n_inst)))[n_inst // 2][1]
    with gcs_amd.Engine(counts[med]) as e1:
        e1.load_generated([specs[med]])
        p1, _ = e1.autotune(gcs_amd.params(G=1.0, eps2=float(np.float32(0.02) ** 2), dt=0.005,
                                           theta=0.5, leaf_size=leaf,
                                           precision=gcs_amd.PREC_COMPENSATED, use_graph=1),
                            reps=2)
    ps = 0
    pt = 0.0
    for sp, n in list(zip(specs, counts))[:plain_sample]:
        e1 = gcs_amd.Engine(n)
        e1.load_generated([sp])
        e1.step(p1, 2)
        t1 = time.perf_counter()
        e1.step(p1, plain_steps, sync=False)
        e1.sync()
        pt += time.perf_counter() - t1
        ps += n * plain_steps
        e1.close()
    plain = ps / pt
    return {"instances": n_inst, "particles": total, "steps": steps,
            "ms_per_step": 1e3 * el / steps, "value": batched,
            "autotuned": {"leaf": p.leaf_size, "block": p.block_size,
                           "waves_per_simd": p.waves_per_simd, "group": p.group_size,
                           "team": p.team},
            "plain_one_at_a_time": plain, "speedup_vs_plain": batched / plain,
            "plain_autotuned": {"leaf": p1.leaf_size, "block": p1.block_size,
                                "waves_per_simd": p1.waves_per_simd, "group": p1.group_size,
                                "team": p1.team, "tuned_on": int(counts[med])},
            "plain_sample": "%d collisions x %d steps" % (min(plain_sample, n_inst), plain_steps)}


def served_run(window, rate, max_resident, leaf, seed=1005, budget=200e9, steps_per_round=25,
               gen_threads=0):
    """Served form: Poisson arrivals at `rate` jobs/s over `window` seconds
    (N log-uniform in [8192, 524288], steps uniform in [200, 2000]), admitted
    under the HBM budget and a resident-particle target, continuously batched.
    The comparator runs the same job list one job at a time (max_jobs = 1)
    over the same window.  Each leg runs at the settings the autotuner picks
    for it: a forest of about 10M particles drawn from the job distribution
    for the served leg, one job of the mean size (124,095) for the
    comparator.  Latency = finish - arrival, over finished jobs.
    gen_threads host threads generate the jobs' initial conditions (0: each
    job is generated on the device at admission)."""
    import numpy as np
    import gcs_amd

    jobs = gcs_amd.poisson_jobs(seed, rate, window, 8192, 524288, 200, 2000)
    eps2 = float(np.float32(0.02) ** 2)
    # Traversal settings for both legs: the autotuner on a forest of about
    # 10M particles drawn from the same job-size distribution.
    specs, tot = [], 0
    while tot < 10_000_000:
        sp = gcs_amd.random_spec(seed + 1, len(specs), 8192, 524288)
        specs.append(sp)
        tot += sp.n_a + sp.n_b
    with gcs_amd.Engine(tot, len(specs)) as e:
        e.load_generated(specs)
        tp, _ = e.autotune(gcs_amd.params(G=1.0, eps2=eps2, dt=0.005, theta=0.5,
                                          leaf_size=leaf, precision=gcs_amd.PREC_COMPENSATED,
                                          use_graph=1), reps=2)
    tuned = dict(leaf_size=tp.leaf_size, group_size=tp.group_size, block_size=tp.block_size,
                 waves_per_simd=tp.waves_per_simd, team=tp.team)
    one = gcs_amd.random_spec(seed + 2, 0, 124095, 124095)
    with gcs_amd.Engine(one.n_a + one.n_b) as e:
        e.load_generated([one])
        t1, _ = e.autotune(gcs_amd.params(G=1.0, eps2=eps2, dt=0.005, theta=0.5,
                                          leaf_size=leaf, precision=gcs_amd.PREC_COMPENSATED,
                                          use_graph=1), reps=2)
    tuned_one = dict(leaf_size=t1.leaf_size, group_size=t1.group_size,
                     block_size=t1.block_size, waves_per_simd=t1.waves_per_simd, team=t1.team)
    out = {"window_s": window, "rate_jobs_per_s": rate, "jobs_offered": len(jobs),
           "offered_particle_steps_per_s": sum(n * k for _, n, k, _ in jobs) / window,
           "hbm_budget_bytes": budget,
           "generation": "device, at admission" if gen_threads <= 0 else
           "%d host threads" % gen_threads,
           "autotuned": {"leaf": tp.leaf_size, "group": tp.group_size, "block": tp.block_size,
                         "waves_per_simd": tp.waves_per_simd, "team": tp.team,
                         "tuning_forest": {"jobs": len(specs), "particles": tot}},
           "autotuned_one_at_a_time": {"leaf": t1.leaf_size, "group": t1.group_size,
                                       "block": t1.block_size,
                                       "waves_per_simd": t1.waves_per_simd, "team": t1.team,
                                       "tuned_on": one.n_a + one.n_b}}
    for name, kw, tn in (("served", dict(max_resident=max_resident), tuned),
                         ("one_at_a_time", dict(max_jobs=1, max_resident=524288), tuned_one)):
        with gcs_amd.Server(hbm_budget_bytes=budget, steps_per_round=steps_per_round,
                            gen_threads=gen_threads, G=1.0, eps2=eps2, dt=0.005, theta=0.5,
                            precision=gcs_amd.PREC_COMPENSATED, use_graph=1, **tn,
                            **kw) as srv:
            srv.submit(jobs, seed)
            srv.run(window)
            st = srv.stats()
            lat = np.array([r["finish"] - r["arrival"] for r in srv.results()])
        out[name] = {
            "value": st["particle_steps"] / st["run_s"], "unit": "particle-steps/s",
            "jobs_finished": int(st["finished"]), "jobs_admitted": int(st["admitted"]),
            "jobs_waiting": int(st["waiting"]), "mean_resident_particles": st["mean_resident"],
            "resident_cap": int(st["resident_cap"]), "rounds": int(st["rounds"]),
            "stepping_fraction": st["stepping_s"] / st["run_s"],
            "steady_value": (st["full_particle_steps"] / st["full_s"]) if st["full_s"] > 0
            else None,
            "steady_s": st["full_s"],
            "p50_latency_s": float(np.quantile(lat, 0.5)) if len(lat) else None,
            "p99_latency_s": float(np.quantile(lat, 0.99)) if len(lat) else None,
        }
        print("served leg %s: %s" % (name, json.dumps(out[name])), file=sys.stderr, flush=True)
    out["speedup_vs_one_at_a_time"] = out["served"]["value"] / out["one_at_a_time"]["value"]
    if out["served"]["steady_value"] and out["one_at_a_time"]["steady_value"]:
        out["steady_speedup_vs_one_at_a_time"] = (out["served"]["steady_value"] /
                                                  out["one_at_a_time"]["steady_value"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="plain-2x131072", choices=sorted(CONFIGS))
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--wps", type=int, default=4, help="traversal waves per SIMD")
    ap.add_argument("--group", type=int, default=96, help="largest traversal group")
    ap.add_argument("--team", type=int, default=4,
                    help="waves per traversal group: 1, or 2, 4, 8 (a block of that many "
                         "waves per group; --block is then 64 x team)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick leaf/block/waves per SIMD with the autotuner instead of "
                         "--leaf/--block/--wps (the batched, served and precision legs always "
                         "autotune)")
    ap.add_argument("--no-autotune", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", type=int, default=256,
                    help="instances of the batched form measured beside the plain step "
                         "(0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=30)
    ap.add_argument("--precision", type=int, default=0,
                    help="steps of the 2 x 1,048,576 precision study (0 = skip)")
    ap.add_argument("--precision-targets", type=int, default=65536)
    ap.add_argument("--served", type=float, default=20.0,
                    help="seconds of the served form (each of two legs; 0 = skip)")
    ap.add_argument("--served-rate", type=float, default=40.0)
    ap.add_argument("--served-gen-threads", type=int, default=0,
                    help="host threads generating the served jobs (0 = on the device at "
                         "admission)")
    ap.add_argument("--served-resident", type=int, default=1 << 24,
                    help="resident-particle admission target of the served form")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU: the library has no multi-GPU path "
              "(independent collisions shard across GPUs without exchange)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import gcs_amd

    per, eps, dt, theta = CONFIGS[a.config]
    stars = gcs_amd.collision(per, 1002)
    n = len(stars)
    eps2 = float(np.float32(eps) ** 2)
    p = gcs_amd.params(G=1.0, eps2=eps2, dt=dt, theta=theta, leaf_size=a.leaf,
                       block_size=a.block, waves_per_simd=a.wps, group_size=a.group, team=a.team,
                       use_graph=1)
    e = gcs_amd.Engine(n)
    e.load(stars)
    if a.autotune and not a.no_autotune:
        p, _ = e.autotune(p, reps=3)
        a.leaf, a.block, a.wps, a.group, a.team = (p.leaf_size, p.block_size,
                                                   p.waves_per_simd, p.group_size, p.team)

    # Accelerations of the initial state (for the accuracy check).
    e.compute_accels(p)
    acc0 = e.accels()
    sorted0 = e.stars()

    e.step(p, a.warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(p, a.steps, sync=False)
    e.sync()
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / a.steps
    value = n * a.steps / el

    # Roofline of the traversal kernel: event-timed eager steps.
    pe = gcs_amd.params(G=1.0, eps2=eps2, dt=dt, theta=theta, leaf_size=a.leaf,
                        block_size=a.block, waves_per_simd=a.wps, group_size=a.group,
                        team=a.team, use_graph=0)
    trav = []
    stages = []
    inter = None
    for _ in range(5):
        e.step(pe, 1)
        st = e.stage_ms()
        stages.append(st)
        trav.append(float(st[5]))
        inter = e.interactions()
    trav_ms = float(np.mean(trav))
    n_inter = inter[0] + inter[1]
    achieved = FLOP_PER_INTERACTION * n_inter / (trav_ms * 1e-3) / 1e12
    executed = ((EXEC_FLOP_NODE * inter[0] + EXEC_FLOP_PARTICLE * inter[1])
                / (trav_ms * 1e-3) / 1e12)
    stage_ms = [float(x) for x in np.mean(np.array(stages), 0)]
    tree_ms = sum(stage_ms[:5])
    tree_gbs = TREE_BYTES_PER_PARTICLE * n / (tree_ms * 1e-3) / 1e9
    info = e.engine_info()
    e.close()

    line = {
        "metric": "particle-steps/s @theta=0.5, err<=1e-3 (plain GPU step)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-galaxy generator)",
        "config": {"workload": a.config, "particles": n, "theta": theta, "eps": eps, "dt": dt,
                   "leaf": a.leaf, "block": a.block, "waves_per_simd": a.wps,
                   "group": a.group, "team": a.team,
                   "autotuned": bool(a.autotune and not a.no_autotune),
                   "multipole": "quadrupole",
                   "parallelism": "single GPU"},
        "interactions_per_step": n_inter,
        "interactions_split": {"node": inter[0], "particle": inter[1]},
        "interactions_per_s": n_inter / (trav_ms * 1e-3),
        "stage_ms": dict(zip(["bounds_morton", "sort", "permute", "build", "aggregate",
                              "traverse", "integrate"], stage_ms)),
        # Compute-bound: priced against the FP32 matrix-core figure (the
        # contract's compute class); the kernel itself issues vector FP32.
        "roofline": {"bound": "mfma", "issue": "vector FP32 (VALU), packed FP32 off",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": TRAVERSE_TRAFFIC.get((a.config, a.leaf, a.block, a.wps, a.group,
                                                      a.team)),
                     "kernel": "k_traverse_team" if a.team > 1 else "k_traverse_group",
                     "kernel_ms": trav_ms,
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "executed": {"tflops": executed,
                                  "flop_per_node": EXEC_FLOP_NODE,
                                  "flop_per_particle": EXEC_FLOP_PARTICLE,
                                  "frac_of_peak": executed / PEAK_FP32_TFLOPS,
                                  "frac_of_unpacked_vector_peak":
                                      executed / PEAK_FP32_VECTOR_TFLOPS}},
        "roofline_tree": {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS,
                          "achieved": tree_gbs, "frac": tree_gbs / PEAK_HBM_GBS,
                          "bytes_per_particle": TREE_BYTES_PER_PARTICLE,
                          "stages": "bounds/morton, sort, permute, build, aggregate",
                          "ms": tree_ms},
        "cpu_baseline": None,
        "engine": info,
    }
    if a.batched > 0:
        line["batched"] = batched_run(a.batched, a.batched_steps, 3, a.leaf)
    if a.precision > 0:
        line["precision"] = precision_study(a.precision, a.precision_targets, a.cpu_seconds)
    if a.served > 0:
        line["served"] = served_run(a.served, a.served_rate, a.served_resident, a.leaf,
                                    gen_threads=a.served_gen_threads)
    if not a.no_cpu:
        cb, acc_check = cpu_baseline(sorted0, p, acc0, a.cpu_seconds)
        line["cpu_baseline"] = cb
        line["accuracy"] = acc_check
        line["speedup_vs_cpu"] = value / cb["value"]
        line["cpu_equal_accuracy"] = cpu_equal_accuracy(sorted0, p, value)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
