#!/usr/bin/env python3
"""Benchmark: particle-steps/s of the per-frame N-body step on one MI355X,
served/batched against plain one-at-a-time runs, at theta = 0.5 with the
acceleration error checked against an FP64 direct sum.

Headline (`value`, N=1): the batched form -- 256 independent collisions, each
of N log-uniform in [4,096, 65,536] (about 5.4M particles), generated on the
device and resident in HBM, advanced together as one forest (one segmented
sort, one forest build and one persistent traversal per step, replayed from a
hipGraph), compensated positions, softening 0.02, dt 0.005.  Exactly
`--steps` steps are timed after `--warmup` untimed ones.  `vs_plain` runs the
same 256 collisions one at a time through the plain path, the same number of
steps each, all of them.

Beside it, in the same line:
  served    the continuous-batching job server on a Poisson stream (40 jobs/s,
            N log-uniform in [8,192, 524,288], 200-2,000 steps) for `--served`
            seconds per leg, against the same job list one job at a time; job
            latency over every job that arrived (unfinished jobs counted with
            their elapsed time, a lower bound);
  plain     one collision of 2 x 131,072 particles, plain FP32;
  c3        one collision of 2 x 1,048,576 particles: step rate, error on
            4,096 FP64 targets, energy drift with plain and compensated
            positions, and the reference CPU octree on the same input.

Far field: softened quadrupoles per accepted node, and for nodes farther than
`--local-kappa` (4) half-diagonals of the target group's box a second-order
local expansion at the box centre (20 coefficients per group, applied per
target), which keeps the error within 1e-3 at theta 0.5 (checked every run).

`roofline` is the dominant kernel of the headline workload, the group
traversal: 20 flop per particle or node interaction, 119 per local-expansion
term and 72 per target's local evaluation (all counted exactly by the kernel)
over its average duration from HIP events on the engine stream, against
MI355X's FP32
vector peak (157.3 TFLOP/s with packed FP32; the build issues unpacked FP32,
whose ceiling is 78.6 -- both fractions are given).  `cpu_baseline` is the
reference's own CPU octree (compiled from its sources into oracle/_ref) timed
on the host cores on a bounded sample of the headline's collisions.

Traversal settings come from the autotuner; its table for this GPU is read
from gcs_amd/autotune_gfx950.txt when present (--retune ignores it).
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "galaxy-collision-simulator_amd"))

METRIC = "particle-steps/s @θ=0.5, err≤1e-3: served/batched vs plain; p50/p99 latency"
PEAK_FP32_TFLOPS = 157.3
PEAK_FP32_VECTOR_TFLOPS = 78.6  # vector FP32 without packed operations
FLOP_PER_INTERACTION = 20
# A node's contribution to a group's local expansion (local_terms) and one
# target's evaluation of the expansion (local_apply), executed FP32 operations
# of csrc/gcs_traverse.hpp, a multiply-add counted as two.
FLOP_PER_LOCAL_TERM = 119
FLOP_PER_LOCAL_APPLY = 72
# FP32 operations the kernel executes per interaction (a multiply-add counted as
# two, the reciprocal square root as one): softened quadrupole node 50,
# particle 19 (interact_n / interact_p in csrc/gcs_traverse.hpp).
EXEC_FLOP_NODE = 50
EXEC_FLOP_PARTICLE = 19
PEAK_HBM_GBS = 8000.0
TREE_BYTES_PER_PARTICLE = 376  # Morton 24 + sort 24 + permute 68 + init 88 + build 24 + agg 148
EPS = 0.02
DT = 0.005
LOCAL_KAPPA = 4.0  # --local-kappa
TUNE_TABLE = os.path.join(REPO, "galaxy-collision-simulator_amd", "gcs_amd",
                          "autotune_gfx950.txt")

# HBM bytes per traversal launch from rocprofv3 FETCH_SIZE (x2.000) and
# WRITE_SIZE (x1.000) passes, scaled by a device copy of known size; valid for
# the (workload, leaf, block, waves per SIMD, group, team, local_kappa)
# measured: 704,515 KiB fetched, 625,298 KiB written per launch.
TRAVERSE_TRAFFIC = {("batched-256", 4, 256, 4, 96, 1, 4.0): 2.083e9}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def settings(p):
    return {"leaf": p.leaf_size, "group": p.group_size, "block": p.block_size,
            "team": p.team, "waves_per_simd": p.waves_per_simd}


def accuracy_check(stars, p, acc_gpu, targets):
    """CPU checker of the GPU accelerations: FP64 direct sum (the oracle) on
    `targets` evenly spaced particles; relative RMS and p99 per-particle
    relative error."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    n = len(stars)
    tgt = np.linspace(0, n - 1, min(targets, n)).astype(np.int64)
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
    cores; particle-steps/s, frames, kind."""
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


def kernel_roofline(e, p, n_steps=5):
    """Event-timed eager steps of the current state: the traversal's mean
    duration (HIP events on the engine stream around its launch), its
    interaction counts (node, particle, far monopole, local-expansion terms,
    node-target pairs those terms stand for), and the tree stage's time."""
    import numpy as np
    import gcs_amd
    pe = gcs_amd.params(**{f: getattr(p, f) for f, _ in p._fields_})
    pe.use_graph = 0
    trav, stages = [], []
    cnt = [0] * 5
    for _ in range(n_steps):
        e.step(pe, 1)
        st = e.stage_ms()
        stages.append(st)
        trav.append(float(st[5]))
        cnt = e.interaction_counts()
    trav_ms = float(np.mean(trav))
    applies = e.count()[1] if p.local_kappa > 0 else 0
    flop = (FLOP_PER_INTERACTION * (cnt[0] + cnt[1] + cnt[2]) + FLOP_PER_LOCAL_TERM * cnt[3]
            + FLOP_PER_LOCAL_APPLY * applies)
    achieved = flop / (trav_ms * 1e-3) / 1e12
    executed = ((EXEC_FLOP_NODE * cnt[0] + EXEC_FLOP_PARTICLE * (cnt[1] + cnt[2])
                 + FLOP_PER_LOCAL_TERM * cnt[3] + FLOP_PER_LOCAL_APPLY * applies)
                / (trav_ms * 1e-3) / 1e12)
    stage_ms = [float(x) for x in np.mean(np.array(stages), 0)]
    return trav_ms, (cnt, applies), achieved, executed, stage_ms


def roofline_obj(workload, p, trav_ms, inter, achieved, executed):
    cnt, applies = inter
    key = (workload, p.leaf_size, p.block_size, p.waves_per_simd, p.group_size, p.team,
           p.local_kappa)
    return {"bound": "fp32-valu",
            "issue": "vector FP32 (VALU) with packed FP32 off; no matrix-core work",
            "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "frac_of_unpacked_vector_peak": achieved / PEAK_FP32_VECTOR_TFLOPS,
            "traffic": TRAVERSE_TRAFFIC.get(key),
            "kernel": "k_traverse_team" if p.team > 1 else "k_traverse_group",
            "kernel_ms": trav_ms, "flop_per_interaction": FLOP_PER_INTERACTION,
            "flop_per_local_term": FLOP_PER_LOCAL_TERM,
            "flop_per_local_apply": FLOP_PER_LOCAL_APPLY,
            "interactions_per_launch": cnt[0] + cnt[1] + cnt[2],
            "interactions_split": {"node": cnt[0], "particle": cnt[1], "far_monopole": cnt[2],
                                   "local_terms": cnt[3], "local_applies": applies,
                                   "node_target_pairs_via_local": cnt[4]},
            "local_kappa": p.local_kappa,
            "executed": {"tflops": executed, "flop_per_node": EXEC_FLOP_NODE,
                         "flop_per_particle": EXEC_FLOP_PARTICLE,
                         "frac_of_unpacked_vector_peak": executed / PEAK_FP32_VECTOR_TFLOPS}}


def tree_roofline(stage_ms, n):
    tree_ms = sum(stage_ms[:5])
    gbs = TREE_BYTES_PER_PARTICLE * n / (tree_ms * 1e-3) / 1e9
    return {"bound": "hbm", "unit": "GB/s", "peak": PEAK_HBM_GBS, "achieved": gbs,
            "frac": gbs / PEAK_HBM_GBS, "bytes_per_particle": TREE_BYTES_PER_PARTICLE,
            "stages": "bounds/morton, sort, permute, build, aggregate", "ms": tree_ms,
            "stage_ms": dict(zip(["bounds_morton", "sort", "permute", "build", "aggregate",
                                  "traverse", "integrate"], stage_ms))}


def base_params(**kw):
    import numpy as np
    import gcs_amd
    kw.setdefault("local_kappa", LOCAL_KAPPA)
    return gcs_amd.params(G=1.0, eps2=float(np.float32(EPS) ** 2), dt=DT, theta=0.5,
                          use_graph=1, **kw)


# ---------------------------------------------------------------------------
# Headline: the batched form (256 collisions) against one at a time
# ---------------------------------------------------------------------------

def batched_leg(n_inst, steps, warmup, seed=1004, check_instances=8, check_targets=512,
                vs_plain=True, eager=False):
    import numpy as np
    import gcs_amd

    specs = [gcs_amd.random_spec(seed, k, 4096, 65536) for k in range(n_inst)]
    counts = [sp.n_a + sp.n_b for sp in specs]
    total = int(sum(counts))
    e = gcs_amd.Engine(total, n_inst)
    e.load_generated(specs)  # initial conditions generated on the device
    p, tune_ms = e.autotune(base_params(precision=gcs_amd.PREC_COMPENSATED), reps=2)
    if eager:
        p.use_graph = 0
    say("batched: %d collisions, %d particles, settings %s" % (n_inst, total, settings(p)))
    # Accuracy of the first frame on a sample of instances (CPU checker).
    e.compute_accels(p)
    acc = e.accels()
    st = e.stars()
    offs = [e.info(i)[1] for i in range(n_inst)]
    pick = np.linspace(0, n_inst - 1, check_instances).astype(int)
    errs = []
    for i in pick:
        sl = slice(offs[i], offs[i] + counts[i])
        errs.append(accuracy_check(st[sl], p, acc[sl], check_targets)["rel_rms_vs_fp64"])
    # Timed region: exactly `steps` replayed forest steps.
    e.step(p, warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(p, steps, sync=False)
    e.sync()
    el = time.perf_counter() - t0
    trav_ms, inter, achieved, executed, stage_ms = kernel_roofline(e, p)
    e.close()
    # One at a time: every collision alone through the plain path (one engine
    # reloaded per collision, graph-replayed steps), at the settings tuned for
    # a single collision of the median size.
    med = sorted(zip(counts, range(n_inst)))[n_inst // 2][1]
    value = total * steps / el
    out = {
        "value": value, "el": el, "ms_per_step": 1e3 * el / steps, "particles": total,
        "instances": n_inst, "counts": counts, "specs": specs, "settings": settings(p),
        "params": p, "tune_ms": tune_ms,
        "accuracy": {"rel_rms_vs_fp64_max": max(errs), "rel_rms_vs_fp64_mean": float(np.mean(errs)),
                     "instances_checked": len(errs), "targets_per_instance": check_targets},
        "vs_plain": None,
        "roofline": roofline_obj("batched-256", p, trav_ms, inter, achieved, executed),
        "tree": tree_roofline(stage_ms, total),
    }
    if not vs_plain:
        return out
    with gcs_amd.Engine(max(counts)) as e1:
        e1.load_generated([specs[med]])
        p1, _ = e1.autotune(base_params(precision=gcs_amd.PREC_COMPENSATED), reps=2)
        pt = 0.0
        for sp in specs:
            e1.load_generated([sp])
            e1.step(p1, warmup)
            e1.sync()
            t1 = time.perf_counter()
            e1.step(p1, steps, sync=False)
            e1.sync()
            pt += time.perf_counter() - t1
    plain = total * steps / pt
    say("batched: %.4g particle-steps/s, one at a time %.4g (%.2fx)" % (value, plain,
                                                                       value / plain))
    out["vs_plain"] = {"plain_one_at_a_time": plain, "speedup": value / plain,
                       "plain_settings": settings(p1), "plain_tuned_on": int(counts[med]),
                       "collisions": n_inst, "steps_each": steps, "plain_s": pt}
    return out


def batched_cpu_baseline(b, sample_seconds):
    """The reference CPU octree frame on the headline's collisions, one after
    another, until about `sample_seconds` of host time."""
    import gcs_amd
    p = base_params()
    done, ps, t = 0, 0, 0.0
    kind = None
    for sp, n in zip(b["specs"], b["counts"]):
        stars = gcs_amd.generate(sp)
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle
        kind = "reference" if oracle.ref_available() else "port"
        step = (lambda s: oracle.ref_octree_step(s, p.G, float(p.eps2), p.dt)) \
            if kind == "reference" \
            else (lambda s: oracle.octree_step(s, p.G, float(p.eps2), p.dt, 0.8, 32))
        if done == 0:
            step(stars)  # warm-up
        t0 = time.perf_counter()
        step(stars)
        t += time.perf_counter() - t0
        ps += n
        done += 1
        if t > sample_seconds:
            break
    return {"value": ps / t, "unit": "particle-steps/s", "cores": threads(), "kind": kind,
            "sample": "one frame of each of the first %d of the headline's %d collisions "
                      "(%d particles) with the reference CPU octree (theta=0.8, leaf 32, as "
                      "compiled from gravity_aggregator.cpp/engine.cpp), OMP_NUM_THREADS=%d, "
                      "%.1f s" % (done, b["instances"], ps, threads(), t)}


# ---------------------------------------------------------------------------
# Single collisions
# ---------------------------------------------------------------------------

def plain_leg(per, steps, warmup, seed, cpu_seconds):
    """One collision of 2 x per particles, plain FP32, graph-replayed steps."""
    import gcs_amd
    stars = gcs_amd.collision(per, seed)
    n = len(stars)
    with gcs_amd.Engine(n) as e:
        e.load(stars)
        p, _ = e.autotune(base_params(), reps=2)
        e.compute_accels(p)
        acc0 = e.accels()
        sorted0 = e.stars()
        e.step(p, warmup)
        e.sync()
        t0 = time.perf_counter()
        e.step(p, steps, sync=False)
        e.sync()
        el = time.perf_counter() - t0
        trav_ms, inter, achieved, executed, stage_ms = kernel_roofline(e, p)
    value = n * steps / el
    out = {"workload": "plain-2x%d" % per, "particles": n, "value": value,
           "unit": "particle-steps/s", "steps": steps, "ms_per_step": 1e3 * el / steps,
           "settings": settings(p),
           "accuracy": accuracy_check(sorted0, p, acc0, 4096),
           "roofline": roofline_obj("plain-2x%d" % per, p, trav_ms, inter, achieved, executed),
           "tree": tree_roofline(stage_ms, n)}
    if cpu_seconds > 0:
        rate, k, kind = cpu_octree_rate(sorted0, p, cpu_seconds)
        out["cpu_octree"] = {"value": rate, "frames": k, "kind": kind, "cores": threads(),
                             "theta": 0.8, "leaf": 32}
        out["speedup_vs_cpu_octree"] = value / rate
        out["cpu_equal_accuracy"] = cpu_equal_accuracy(sorted0, p, value, acc0)
    say("plain-2x%d: %.4g particle-steps/s" % (per, value))
    return out


def cpu_equal_accuracy(stars, p, gpu_value, acc_gpu, thetas=(0.6, 0.5, 0.4, 0.35, 0.3)):
    """The CPU octree (the oracle's restatement of the reference octree, with
    the opening angle as a parameter, leaf cap 32, host threads) at the first
    angle whose accelerations meet 1e-3 relative RMS against the same FP64
    check as the GPU: that angle, its error, one frame's time, the GPU's
    speed-up over it, and the RMS difference between the GPU's and that
    octree's accelerations; and the first angle at which that difference
    is within 1e-3 (`both_within_1e-3_at_theta`: the GPU then meets 1e-3
    against both the FP64 sum and the octree).  The reference hard-codes
    theta = 0.8, which misses 1e-3 by an order of magnitude."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    n = len(stars)
    tried = []
    out, both = None, None
    for th in (0.8,) + tuple(thetas):
        t0 = time.perf_counter()
        s1, acc, _ = oracle.octree_step(stars, p.G, float(p.eps2), p.dt, th, 32,
                                        integrate=False, want_acc=True)
        el = time.perf_counter() - t0
        chk = accuracy_check(s1, p, acc, 4096)
        # The octree re-sorts its input by its own Morton keys: match its
        # particles to the GPU's by position.
        ia = np.lexsort((stars["z"], stars["y"], stars["x"]))
        ib = np.lexsort((s1["z"], s1["y"], s1["x"]))
        ag = acc_gpu[ia, :3].astype(np.float64)
        ao = acc[ib].astype(np.float64)
        gpu_vs_oct = float(np.sqrt(((ag - ao) ** 2).sum() / (ao ** 2).sum()))
        tried.append({"theta": th, "rel_rms_vs_fp64": chk["rel_rms_vs_fp64"],
                      "gpu_vs_octree_rel_rms": gpu_vs_oct,
                      "frame_s": el, "particle_steps_per_s": n / el})
        if out is None and th < 0.8 and chk["rel_rms_vs_fp64"] <= 1e-3:
            out = {"theta": th, "leaf": 32, "kind": "port",
                   "rel_rms_vs_fp64": chk["rel_rms_vs_fp64"],
                   "gpu_vs_octree_rel_rms": gpu_vs_oct,
                   "value": n / el, "unit": "particle-steps/s", "cores": threads(),
                   "speedup_vs_cpu": gpu_value / (n / el)}
        if gpu_vs_oct <= 1e-3:
            both = th
        if out is not None and both is not None:
            break
    out = out or {"theta": None}
    out["both_within_1e-3_at_theta"] = both
    out["tried"] = tried
    return out


def c3_leg(steps, warmup, drift_steps, cpu_seconds, targets, seed=1003):
    """One collision of 2 x 1,048,576 particles at theta 0.5: plain step rate
    and the reference CPU octree on the same input, the error on `targets`
    FP64 direct-sum targets, and the relative energy drift over drift_steps
    steps with plain and with compensated positions (FP64 diagnostics)."""
    import gcs_amd
    stars = gcs_amd.collision(1048576, seed)
    n = len(stars)
    out = {"workload": "2x1048576", "particles": n, "theta": 0.5}
    with gcs_amd.Engine(n) as e:
        e.load(stars)
        p, _ = e.autotune(base_params(), reps=2)
        e.compute_accels(p)
        acc0 = e.accels()
        sorted0 = e.stars()
        e.step(p, warmup)
        e.sync()
        t0 = time.perf_counter()
        e.step(p, steps, sync=False)
        e.sync()
        el = time.perf_counter() - t0
        trav_ms, inter, achieved, executed, stage_ms = kernel_roofline(e, p, 3)
    out.update({"value": n * steps / el, "unit": "particle-steps/s", "steps": steps,
                "ms_per_step": 1e3 * el / steps, "settings": settings(p),
                "roofline": roofline_obj("2x1048576", p, trav_ms, inter, achieved, executed),
                "tree": tree_roofline(stage_ms, n)})
    out["accuracy"] = accuracy_check(sorted0, p, acc0, targets)
    for prec, name in ((gcs_amd.PREC_PLAIN, "plain"), (gcs_amd.PREC_COMPENSATED, "compensated")):
        if drift_steps < 2:
            break
        kw = {f: getattr(p, f) for f, _ in p._fields_}
        kw["precision"] = prec
        with gcs_amd.Engine(n) as e:
            e.load(stars)
            kw["diagnostics"] = 1
            e.step(gcs_amd.params(**kw), 1)
            d0 = e.diagnostics()
            kw["diagnostics"] = 0
            e.step(gcs_amd.params(**kw), drift_steps - 2)
            kw["diagnostics"] = 1
            e.step(gcs_amd.params(**kw), 1)
            d1 = e.diagnostics()
        e0, e1 = d0[0] + d0[1], d1[0] + d1[1]
        out["energy_drift_" + name] = {"steps": drift_steps, "relative": (e1 - e0) / abs(e0),
                                       "momentum_change": float(
                                           sum((d1[k] - d0[k]) ** 2 for k in (2, 3, 4)) ** 0.5)}
    if cpu_seconds > 0:
        rate, k, kind = cpu_octree_rate(sorted0, p, cpu_seconds, max_steps=5)
        out["cpu_octree"] = {"value": rate, "frames": k, "kind": kind, "cores": threads(),
                             "theta": 0.8, "leaf": 32}
        out["speedup_vs_cpu_octree"] = out["value"] / rate
    say("c3: %.4g particle-steps/s, rms %.3g" % (out["value"],
                                                  out["accuracy"]["rel_rms_vs_fp64"]))
    return out


def precision_study(steps, targets, seed=1003):
    """The full precision form (--precision STEPS): 2 x 1,048,576 at theta 0.5
    and 0.7, errors on `targets` FP64 targets, energy drift over `steps` steps
    with plain and compensated positions."""
    import numpy as np
    import gcs_amd
    stars = gcs_amd.collision(1048576, seed)
    n = len(stars)
    out = {"particles": n, "steps": steps, "runs": []}
    for theta in (0.5, 0.7):
        base = base_params(theta=theta)
        with gcs_amd.Engine(n) as e:
            e.load(stars)
            base, _ = e.autotune(base, reps=2)
            e.compute_accels(base)
            acc = e.accels()
            s0 = e.stars()
        run = {"theta": theta, "settings": settings(base)}
        run.update(accuracy_check(s0, base, acc, targets))
        for prec, name in ((gcs_amd.PREC_PLAIN, "plain"),
                           (gcs_amd.PREC_COMPENSATED, "compensated")):
            kw = {f: getattr(base, f) for f, _ in base._fields_}
            kw["precision"] = prec
            with gcs_amd.Engine(n) as e:
                e.load(stars)
                kw["diagnostics"] = 1
                e.step(gcs_amd.params(**kw), 1)
                d0 = e.diagnostics()
                kw["diagnostics"] = 0
                t0 = time.perf_counter()
                e.step(gcs_amd.params(**kw), steps - 2)
                el = time.perf_counter() - t0
                kw["diagnostics"] = 1
                e.step(gcs_amd.params(**kw), 1)
                d1 = e.diagnostics()
            e0, e1 = d0[0] + d0[1], d1[0] + d1[1]
            run[name] = {"energy_drift_rel": (e1 - e0) / abs(e0),
                         "momentum_change": float(np.linalg.norm(d1[2:5] - d0[2:5])),
                         "ms_per_step": 1e3 * el / (steps - 2)}
        out["runs"].append(run)
        say("precision theta %.1f: %s" % (theta, json.dumps(run)))
    return out


# ---------------------------------------------------------------------------
# Served form
# ---------------------------------------------------------------------------

def latency_summary(jobs, results, window):
    """p50/p99 of job latency (finish - arrival) over every job that arrived
    in the window: finished jobs give their latency, unfinished ones their
    elapsed time at the end of the window (a lower bound), so the quantiles
    are lower bounds whenever any job is unfinished.  Finished-only quantiles
    are given too (biased towards short jobs)."""
    import numpy as np
    done = {r["id"]: r["finish"] - r["arrival"] for r in results}
    arrived = [(jid, t) for jid, _, _, t in jobs if t < window]
    lat = np.array([done.get(jid, window - t) for jid, t in arrived])
    fin = np.array(list(done.values()))
    out = {"jobs_arrived": len(arrived), "jobs_finished": len(done),
           "jobs_censored": len(arrived) - len(done),
           "p50_latency_s": float(np.quantile(lat, 0.5)) if len(lat) else None,
           "p99_latency_s": float(np.quantile(lat, 0.99)) if len(lat) else None,
           "latency_is_lower_bound": len(arrived) > len(done),
           "p50_latency_finished_s": float(np.quantile(fin, 0.5)) if len(fin) else None,
           "p99_latency_finished_s": float(np.quantile(fin, 0.99)) if len(fin) else None}
    return out


def served_leg(window, rate, max_resident, seed=1005, budget=200e9, steps_per_round=25,
               legs=("served", "one_at_a_time"), slice_s=20.0):
    """Each leg runs the job stream for `window` seconds of server time, in
    slices of `slice_s` seconds (the server clock runs on across slices) with
    a progress line after each."""
    import gcs_amd

    jobs = gcs_amd.poisson_jobs(seed, rate, window, 8192, 524288, 200, 2000)
    # Settings for the served leg: the autotuner on a forest of about 10M
    # particles drawn from the job-size distribution; for the comparator, on
    # one job of the mean size (124,095).
    specs, tot = [], 0
    while tot < 10_000_000:
        sp = gcs_amd.random_spec(seed + 1, len(specs), 8192, 524288)
        specs.append(sp)
        tot += sp.n_a + sp.n_b
    with gcs_amd.Engine(tot, len(specs)) as e:
        e.load_generated(specs)
        tp, _ = e.autotune(base_params(precision=gcs_amd.PREC_COMPENSATED), reps=2)
    one = gcs_amd.random_spec(seed + 2, 0, 124095, 124095)
    with gcs_amd.Engine(one.n_a + one.n_b) as e:
        e.load_generated([one])
        t1, _ = e.autotune(base_params(precision=gcs_amd.PREC_COMPENSATED), reps=2)
    out = {"window_s": window, "rate_jobs_per_s": rate, "jobs_offered": len(jobs),
           "offered_particle_steps_per_s": sum(n * k for _, n, k, _ in jobs) / window,
           "hbm_budget_bytes": budget, "generation": "device, at admission",
           "settings": settings(tp), "settings_one_at_a_time": settings(t1)}
    for name, kw, tn in (("served", dict(max_resident=max_resident), tp),
                         ("one_at_a_time", dict(max_jobs=1, max_resident=524288), t1)):
        if name not in legs:
            continue
        tune = {"leaf_size": tn.leaf_size, "group_size": tn.group_size,
                "block_size": tn.block_size, "waves_per_simd": tn.waves_per_simd,
                "team": tn.team}
        with gcs_amd.Server(hbm_budget_bytes=budget, steps_per_round=steps_per_round,
                            gen_threads=0, G=1.0, eps2=tn.eps2, dt=DT, theta=0.5,
                            precision=gcs_amd.PREC_COMPENSATED, use_graph=1, **tune,
                            **kw) as srv:
            srv.submit(jobs, seed)
            done = 0.0
            while done < window:
                part = min(slice_s, window - done)
                srv.run(part)
                done += part
                st = srv.stats()
                say("served leg %s: %.0f / %.0f s, %d jobs finished, %.4g particle-steps/s"
                    % (name, done, window, int(st["finished"]),
                       st["particle_steps"] / max(st["run_s"], 1e-9)))
            st = srv.stats()
            res = srv.results()
        leg = {"value": st["particle_steps"] / st["run_s"], "unit": "particle-steps/s",
               "jobs_admitted": int(st["admitted"]), "jobs_waiting": int(st["waiting"]),
               "mean_resident_particles": st["mean_resident"],
               "resident_cap": int(st["resident_cap"]), "rounds": int(st["rounds"]),
               "stepping_fraction": st["stepping_s"] / st["run_s"],
               "steady_value": (st["full_particle_steps"] / st["full_s"]) if st["full_s"] > 0
               else None, "steady_s": st["full_s"]}
        leg.update(latency_summary(jobs, res, window))
        out[name] = leg
        say("served leg %s: %s" % (name, json.dumps(leg)))
    if "served" not in out or "one_at_a_time" not in out:
        return out
    out["speedup_vs_one_at_a_time"] = out["served"]["value"] / out["one_at_a_time"]["value"]
    if out["served"]["steady_value"] and out["one_at_a_time"]["steady_value"]:
        out["steady_speedup_vs_one_at_a_time"] = (out["served"]["steady_value"] /
                                                  out["one_at_a_time"]["steady_value"])
    return out


def bake_leg(frames, per=150000):
    """The Bake frame at the reference README's conditions (300,000 particles,
    eps^2 = 10, dt = 0.01, the reference's traversal rule; README.md:92-104
    quotes ~31 ms per frame on an RTX 3050 Laptop): the gcs_bake driver's
    MI355X pipeline and the reference's own frame through the cuda_engine.h
    functions, each writing `frames` frames of simulation.bin to a scratch
    file.  Context only: the initial conditions are this repository's
    generator, not the reference's."""
    import subprocess
    import tempfile
    import gcs_amd
    exe = os.path.join(os.path.dirname(gcs_amd.LIB_PATH), "gcs_bake")
    out = {"particles": 2 * per, "frames": frames, "eps2": 10.0, "dt": 0.01,
           "reference_published_ms_per_frame": 31.4,
           "reference_hardware": "RTX 3050 Laptop (README.md:102)"}
    with tempfile.TemporaryDirectory() as d:
        for name, extra in (("pipeline", []), ("via_shim", ["--via-shim"])):
            path = os.path.join(d, "simulation.bin")
            r = subprocess.run([exe, str(frames), "--n", str(per), "--eps2", "10", "--dt",
                                "0.01", "--seed", "1001", "--out", path] + extra,
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                               timeout=300)
            if r.returncode != 0:
                raise RuntimeError("gcs_bake failed: " + r.stderr[-500:])
            out[name] = json.loads(r.stdout.strip().splitlines()[-1])
            out[name]["file_bytes"] = os.path.getsize(path)
            os.remove(path)
    say("bake: pipeline %.3f ms/frame, via shim %.3f ms/frame"
        % (out["pipeline"]["ms_per_frame"], out["via_shim"]["ms_per_frame"]))
    return out


def copy_calibration(mib):
    """Device-to-device copies of `mib` MiB through the HIP runtime (the
    runtime's copy kernel), for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import ctypes
    import gcs_amd
    gcs_amd.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    size = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(a), size) == 0 and h.hipMalloc(ctypes.byref(b), size) == 0
    assert h.hipMemset(a, 1, size) == 0
    for _ in range(3):
        assert h.hipMemcpy(b, a, size, 3) == 0  # hipMemcpyDeviceToDevice
    assert h.hipDeviceSynchronize() == 0
    h.hipFree(a)
    h.hipFree(b)
    print(json.dumps({"copy_calibration_mib": mib, "copies": 3}), flush=True)


def main():
    global LOCAL_KAPPA
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of the headline")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="batched-256", choices=["batched-256"],
                    help="headline workload (the batched form of 256 collisions)")
    ap.add_argument("--batched", type=int, default=256,
                    help="collisions of the batched headline (0 = skip it: profiling of the "
                         "plain and c3 legs alone)")
    ap.add_argument("--vs-plain", type=int, default=1,
                    help="1: run the headline's collisions one at a time too (0 = skip)")
    ap.add_argument("--eager", action="store_true",
                    help="launch the headline's steps one by one instead of replaying a "
                         "hipGraph (the same kernels; for profilers)")
    ap.add_argument("--served", type=float, default=30.0,
                    help="seconds of each served leg (0 = skip)")
    ap.add_argument("--served-rate", type=float, default=40.0)
    ap.add_argument("--served-resident", type=int, default=1 << 24,
                    help="resident-particle admission target of the served form")
    ap.add_argument("--served-legs", default="served,one_at_a_time",
                    help="legs of the served form to run (comma-separated)")
    ap.add_argument("--plain", type=int, default=131072,
                    help="particles per galaxy of the plain single-collision leg (0 = skip)")
    ap.add_argument("--c3-steps", type=int, default=10,
                    help="timed steps of the 2 x 1,048,576 leg (0 = skip)")
    ap.add_argument("--c3-drift-steps", type=int, default=200)
    ap.add_argument("--precision", type=int, default=0,
                    help="steps of the full 2 x 1,048,576 precision study at theta 0.5 and "
                         "0.7 (0 = skip)")
    ap.add_argument("--precision-targets", type=int, default=65536)
    ap.add_argument("--bake-frames", type=int, default=30,
                    help="frames of each Bake mode at the README's conditions (0 = skip)")
    ap.add_argument("--local-kappa", type=float, default=LOCAL_KAPPA,
                    help="local expansions for nodes beyond this many group half-diagonals "
                         "(0 = off: quadrupoles only)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--retune", action="store_true",
                    help="ignore the committed autotuner table and tune every workload")
    ap.add_argument("--tune-out", default=None, help="write the autotuner table here")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU: the library has no multi-GPU path "
              "(independent collisions shard across GPUs without exchange)",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration)
        return

    LOCAL_KAPPA = a.local_kappa
    import gcs_amd
    if not a.retune and os.path.exists(TUNE_TABLE):
        gcs_amd.autotune_cache_load(TUNE_TABLE)

    if a.batched == 0:
        # Profiling form: only the legs asked for, no headline value.
        line = {"metric": METRIC, "value": None, "unit": "particle-steps/s", "n_gpus": 1,
                "note": "--batched 0: headline skipped (profiling of the other legs)"}
        if a.plain > 0:
            line["plain"] = plain_leg(a.plain, a.steps, a.warmup, 1002, 0)
        if a.c3_steps > 0:
            line["c3"] = c3_leg(a.c3_steps, 2, a.c3_drift_steps, 0, 4096)
        print(json.dumps(line), flush=True)
        return
    b = batched_leg(a.batched, a.steps, a.warmup, vs_plain=bool(a.vs_plain), eager=a.eager)
    line = {
        "metric": METRIC,
        "value": b["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": b["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded two-galaxy generator, generated on the device)",
        "config": {"workload": "batched-256: %d collisions, N log-uniform in [4096, 65536]"
                               % b["instances"],
                   "instances": b["instances"], "particles": b["particles"], "theta": 0.5,
                   "eps": EPS, "dt": DT, "precision": "compensated FP32 hi/lo positions",
                   "multipole": "quadrupole", "local_kappa": LOCAL_KAPPA,
                   "parallelism": "single GPU", **b["settings"]},
        "vs_plain": b["vs_plain"],
        "accuracy": b["accuracy"],
        "roofline": b["roofline"],
        "roofline_tree": b["tree"],
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        line["cpu_baseline"] = batched_cpu_baseline(b, a.cpu_seconds)
        line["speedup_vs_cpu"] = b["value"] / line["cpu_baseline"]["value"]
    if a.served > 0:
        line["served"] = served_leg(a.served, a.served_rate, a.served_resident,
                                    legs=tuple(a.served_legs.split(",")))
    if a.plain > 0:
        line["plain"] = plain_leg(a.plain, a.steps, a.warmup, 1002,
                                  0 if a.no_cpu else a.cpu_seconds / 2)
    if a.c3_steps > 0:
        line["c3"] = c3_leg(a.c3_steps, 2, a.c3_drift_steps, 0 if a.no_cpu else a.cpu_seconds,
                            4096)
    if a.precision > 0:
        line["precision"] = precision_study(a.precision, a.precision_targets)
    if a.bake_frames > 0:
        line["bake"] = bake_leg(a.bake_frames)
    if a.tune_out:
        gcs_amd.autotune_cache_save(a.tune_out)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
