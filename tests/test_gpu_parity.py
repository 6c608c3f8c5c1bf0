"""HIP path vs the CPU oracle, through the C ABI.

Bit-exact: Morton keys, the sort (keys and permutation), the Karras topology,
node aggregates (COM, mass, AABB, width), the kick-drift and colours.
Tolerance (written per test): accelerations -- 2e-5 relative RMS against the
oracle's FP64 evaluation of the same interaction lists, <= 1e-3 relative RMS
against the FP64 direct sum at theta = 0.5.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

G = 1.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def cloud(gcs, n, seed, spread=1.0):
    r = np.random.default_rng(seed)
    st = np.zeros(n, gcs.STAR)
    for k in ("x", "y", "z"):
        st[k] = (r.normal(size=n) * spread).astype(np.float32)
    for k in ("vx", "vy", "vz"):
        st[k] = r.normal(size=n).astype(np.float32) * 0.1
    st["mass"] = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    st["type"] = r.integers(0, 3, n)
    return st


# ---- stand-alone stages --------------------------------------------------------

def test_morton_keys_match_oracle(gcs, orc):
    r = np.random.default_rng(1)
    mn, mx = -7.25, 3.5
    x = np.concatenate([r.uniform(-9, 6, 100000), [mn, mx, np.nan, np.inf, -np.inf]])
    y = r.uniform(-9, 6, len(x))
    z = r.uniform(-9, 6, len(x))
    x, y, z = (a.astype(np.float32) for a in (x, y, z))
    say("morton n=%d" % len(x))
    assert np.array_equal(gcs.morton_keys(x, y, z, mn, mx), orc.morton_n(x, y, z, mn, mx))
    # degenerate span: scale = +inf; points on the span give 0, others saturate
    assert np.array_equal(gcs.morton_keys(x[:100], y[:100], z[:100], 2.0, 2.0),
                          orc.morton_n(x[:100], y[:100], z[:100], 2.0, 2.0))
    two = np.full(50, 2.0, np.float32)
    assert np.all(gcs.morton_keys(two, two, two, 2.0, 2.0) == 0)


def test_morton_keys_match_reference_golden(gcs):
    """HIP Morton keys == the reference's compiled get_morton_code (fixture)."""
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             "morton_ref.npz"))
    for i in range(int(g["cases"])):
        mn, mx = (float(v) for v in g["bounds%d" % i])
        say("morton golden case %d" % i)
        k = gcs.morton_keys(g["x%d" % i], g["y%d" % i], g["z%d" % i], mn, mx)
        assert np.array_equal(k, g["keys%d" % i]), i


@pytest.mark.parametrize("n", [1, 2, 63, 1023, 1024, 1025, 4097, 100003])
def test_sort_pairs_match_oracle(gcs, orc, n):
    r = np.random.default_rng(n)
    keys = r.integers(0, 1 << 63, n, dtype=np.uint64)
    keys[: n // 3] = keys[: n // 3] % np.uint64(13)
    keys[n // 3: n // 2] = np.uint64(0x7FFFFFFFFFFFFFFF)
    say("sort n=%d" % n)
    ko, po = gcs.sort_pairs(keys)
    eo, ep = orc.stable_sort_pairs(keys)
    assert np.array_equal(ko, eo) and np.array_equal(po, ep)


@pytest.mark.parametrize("n", [786431, 786432, 786432 + 4097])
def test_sort_pairs_at_the_tile_size_threshold(gcs, orc, n):
    """Sorts of fewer than 786,432 keys run in 1,024-key blocks, larger ones in
    4,096-key blocks (a last block of one key included): both bit-exact."""
    r = np.random.default_rng(n)
    keys = r.integers(0, 1 << 63, n, dtype=np.uint64)
    keys[::7] = keys[::7] % np.uint64(1000)  # many equal keys
    say("sort n=%d" % n)
    ko, po = gcs.sort_pairs(keys)
    eo, ep = orc.stable_sort_pairs(keys)
    assert np.array_equal(ko, eo) and np.array_equal(po, ep)


def test_sort_pairs_full_64_bits_and_presorted(gcs, orc):
    r = np.random.default_rng(5)
    keys = r.integers(0, 1 << 64, 2_000_000, dtype=np.uint64)  # top bit used too
    for k in (keys, np.sort(keys), np.sort(keys)[::-1].copy()):
        say("sort 2M")
        ko, po = gcs.sort_pairs(k)
        eo, ep = orc.stable_sort_pairs(k)
        assert np.array_equal(ko, eo) and np.array_equal(po, ep)


def karras_cases():
    r = np.random.default_rng(3)
    out = []
    for n in (1, 2, 3, 5, 64, 1000, 5000):
        out.append(np.sort(r.integers(0, 1 << 63, n, dtype=np.uint64)))
        out.append(np.sort(r.integers(0, 4, n, dtype=np.uint64)))
    out.append(np.zeros(3000, np.uint64))
    out.append(np.sort(r.integers(0, 1 << 20, 262144, dtype=np.uint64)))
    return out


@pytest.mark.parametrize("case", range(16))
def test_karras_tree_matches_oracle(gcs, orc, case):
    keys = karras_cases()[case]
    say("karras case %d n=%d" % (case, len(keys)))
    l, r, p = gcs.karras_tree(keys)
    el, er, ep = orc.karras_build(keys)
    assert np.array_equal(l, el) and np.array_equal(r, er) and np.array_equal(p, ep)


# ---- engine: one instance ----------------------------------------------------------

def check_pipeline(gcs, orc, e, stars, p, iid=0, theta=0.5, bounds=None):
    raw, srt, perm = e.keys()
    cpu = orc.build_tree_pipeline(stars, theta, *(bounds or (None, None)),
                                  offset=bool(p.open_offset))
    if bounds is None:
        mn, mx = e.bounds(iid)
        assert (mn, mx) == (np.float32(cpu["mn"]), np.float32(cpu["mx"]))
    assert np.array_equal(raw, cpu["keys"])
    assert np.array_equal(srt, cpu["skeys"])
    assert np.array_equal(perm, cpu["perm"])
    t = e.tree(iid)
    for k in ("left", "right", "parent"):
        assert np.array_equal(t[k], cpu[k]), k
    agg = cpu["agg"]
    for k in ("com4", "width", "bmin", "bmax"):
        assert np.array_equal(t[k].view(np.uint32), agg[k].view(np.uint32)), k
    return cpu


@pytest.mark.parametrize("gsize,offset,team,mono,kappa", [
    (64, 0, 1, 0.0, 0.0), (24, 0, 1, 0.0, 0.0), (128, 0, 1, 0.0, 0.0), (192, 0, 1, 0.0, 0.0),
    (64, 1, 1, 0.0, 0.0), (64, 0, 2, 0.0, 0.0), (24, 0, 2, 0.0, 0.0), (192, 0, 2, 0.0, 0.0),
    (96, 0, 4, 0.0, 0.0), (5, 0, 4, 0.0, 0.0), (192, 0, 8, 0.0, 0.0), (128, 0, 1, 1.7, 0.0),
    (96, 0, 4, 1.7, 0.0), (24, 0, 2, 1.3, 0.0), (64, 1, 1, 2.0, 0.0),
    (128, 0, 1, 0.0, 6.0), (96, 0, 4, 0.0, 6.0), (5, 0, 4, 0.0, 3.0), (192, 0, 8, 0.0, 5.0),
    (64, 0, 1, 1.5, 4.0), (24, 0, 2, 0.0, 6.0)])
def test_engine_c1_group_parity(gcs, orc, gsize, offset, team, mono, kappa):
    stars = gcs.collision(8192, 1001)
    n = len(stars)
    eps2 = np.float32(0.05) ** 2
    p = gcs.params(G=G, eps2=float(eps2), dt=0.01, theta=0.5, leaf_size=16, diagnostics=1,
                   group_size=gsize, open_offset=offset, team=team, mono_far=mono,
                   local_kappa=kappa, block_size=64 * team if team > 1 else 256)
    with gcs.Engine(n) as e:
        e.load(stars)
        say("engine C1 compute, groups of at most %d, opening offset %d, team %d, monopoles "
            "beyond %.1f x the opening radius, local expansions beyond %.1f group radii"
            % (gsize, offset, team, mono, kappa))
        e.compute_accels(p)
        cpu = check_pipeline(gcs, orc, e, stars, p)
        acc = e.accels()
        s = cpu["sorted"]
        grp = orc.tree_groups(cpu["left"], cpu["right"], cpu["first"], cpu["last"], gsize)
        want, msum, cnt, aabs = orc.traverse_group(cpu["left"], cpu["right"], cpu["first"],
                                                   cpu["last"], s["x"], s["y"], s["z"],
                                                   s["mass"], cpu["agg"], G, float(eps2), 64, 16,
                                                   want_abs=True, groups=grp, mono_far=mono,
                                                   local_kappa=kappa)
        assert e.engine_info()["groups"] == len(grp)
        assert e.interaction_counts() == [int(c) for c in cnt]
        assert (cnt[2] > 0) == (mono > 0) and (cnt[3] > 0) == (kappa > 0)
        # Same interaction lists: identical counts.  Same sums: the FP32
        # accumulation error is bounded by a multiple of unit roundoff times
        # the sum of term magnitudes (3e-5 ~ 500 u here; a dropped or doubled
        # term of the lists is larger than this for any particle).
        nn, npart = e.interactions()
        assert (nn, npart) == (int(cnt[0]), int(cnt[1]))
        dev = np.abs(acc[:, :3] - want[:, :3]).max(1) / aabs
        say("C1 vs oracle lists: rel rms %.3g, max dev/sum|terms| %.3g"
            % (rel_rms(acc[:, :3], want[:, :3]), dev.max()))
        assert dev.max() < 3e-5
        assert rel_rms(acc[:, 3], want[:, 3]) < 2e-5
        ref = orc.direct_sum_f64(s["x"], s["y"], s["z"], s["mass"], G, float(eps2))
        err = rel_rms(acc[:, :3], ref[:, :3])
        say("C1 rel rms vs FP64 direct sum: %.3g" % err)
        # The second-order local expansion's truncation error scales as
        # kappa^-3: kappa = 3 (chosen to exercise many local interactions)
        # is outside the 1e-3 accuracy target, kappa >= 4 inside it.
        assert err < (1e-3 if kappa == 0 or kappa >= 4 else 4e-3)
        # potential: second order for quadrupoles, first for far monopoles
        assert rel_rms(acc[:, 3], ref[:, 3]) < (1e-4 if mono == 0 and kappa == 0 else 5e-4)


@pytest.mark.parametrize("team,leaf,kappa", [(8, 64, 0.0), (4, 64, 4.0), (2, 48, 0.0),
                                            (1, 64, 4.0)])
def test_group_traversal_large_leaf(gcs, orc, team, leaf, kappa):
    """Leaf sizes up to the largest allowed: a round of a team of 8 waves can
    add 8 x 64 x 66 particle-list entries, which its lists hold."""
    stars = gcs.collision(8192, 1001)
    eps2 = np.float32(0.05) ** 2
    p = gcs.params(G=G, eps2=float(eps2), dt=0.01, theta=0.5, leaf_size=leaf, group_size=128,
                   team=team, local_kappa=kappa, block_size=64 * team if team > 1 else 64)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        say("team %d, leaf %d, kappa %.0f" % (team, leaf, kappa))
        e.compute_accels(p)
        cpu = check_pipeline(gcs, orc, e, stars, p)
        acc = e.accels()
        s = cpu["sorted"]
        grp = orc.tree_groups(cpu["left"], cpu["right"], cpu["first"], cpu["last"], 128)
        want, msum, cnt, aabs = orc.traverse_group(cpu["left"], cpu["right"], cpu["first"],
                                                   cpu["last"], s["x"], s["y"], s["z"],
                                                   s["mass"], cpu["agg"], G, float(eps2), 128,
                                                   leaf, want_abs=True, groups=grp,
                                                   local_kappa=kappa)
        assert e.interaction_counts() == [int(c) for c in cnt]
        dev = np.abs(acc[:, :3] - want[:, :3]).max(1) / aabs
        say("max dev/sum|terms| %.3g" % dev.max())
        assert dev.max() < 3e-5
        ref = orc.direct_sum_f64(s["x"], s["y"], s["z"], s["mass"], G, float(eps2))
        assert rel_rms(acc[:, :3], ref[:, :3]) < 1e-3


def top_classes(cpu, atile=1024):
    """Nodes of more than `atile` particles (the ones the instance pass
    finishes) per size class (eight per octave), from the oracle's tree."""
    size = (cpu["last"] - cpu["first"] + 1).astype(np.int64)
    size = size[size > atile]
    lg = np.floor(np.log2(size)).astype(np.int64)
    cls = 8 * (lg - 10) + ((size >> (lg - 3)) & 7)
    return len(size), np.bincount(cls)


@pytest.mark.parametrize("case,levels", [("phases", 1), ("global", 1), ("phases", 2),
                                         ("global", 2)])
def test_instance_pass_large_instances(gcs, orc, case, levels):
    """The instance pass on instances with more than 2048 nodes above the
    aggregation subtrees.  In one level: in LDS phases by size class (two
    collided galaxies of 1,048,576), and by climbing with arrival marks when
    one size class alone holds more than 2048 nodes (4.4M particles uniform
    in a cube).  In two levels: the subtrees of fewer than 32,768 particles one
    block each, then the rest.  Pipeline bit-exact against the oracle,
    aggregates included."""
    if case == "phases":
        stars = gcs.collision(1048576, 1007)
    else:
        stars = cloud(gcs, 4404019, 8)
        for k in ("x", "y", "z"):
            stars[k] = np.random.default_rng({"x": 1, "y": 2, "z": 3}[k]).uniform(
                -1, 1, len(stars)).astype(np.float32)
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8, group_size=128,
                   instance_pass=levels)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        say("instance pass, %s, %d level(s): n=%d" % (case, levels, len(stars)))
        e.compute_accels(p)
        cpu = check_pipeline(gcs, orc, e, stars, p)
    nt, hist = top_classes(cpu)
    say("instance pass, %s: %d nodes above the subtrees, largest size class %d"
        % (case, nt, hist.max()))
    assert nt > 2048
    assert (hist.max() <= 2048) == (case == "phases")


def test_engine_reference_mode_parity(gcs, orc):
    stars = cloud(gcs, 20000, 7)
    eps2 = 0.0625
    mn, mx = -6.0, 6.5
    p = gcs.reference_params(G, eps2, 0.01, mn, mx)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        say("reference mode compute")
        e.compute_accels(p)
        cpu = check_pipeline(gcs, orc, e, stars, p, theta=1.0, bounds=(mn, mx))
        s = cpu["sorted"]
        acc = e.accels()
        want, _, cnt = orc.traverse_reference(cpu["left"], cpu["right"], s["x"], s["y"], s["z"],
                                              cpu["agg"]["com4"], cpu["agg"]["width"], G, eps2)
        assert rel_rms(acc[:, :3], want[:, :3]) < 2e-5
        assert e.interactions()[1] == cnt


def test_integrate_and_colours_bit_exact(gcs, orc):
    stars = cloud(gcs, 5000, 8)
    stars["vx"] *= 2000.0  # speeds across the colour ramp
    p = gcs.reference_params(G, 0.0625, 0.01, -8.0, 8.0, use_graph=0)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        e.compute_accels(p)
        before = e.stars()
        acc = e.accels()
        say("integrate step")
        e.step(p, 1)  # same order (stable sort of unchanged keys), same forces
        after = e.stars()
    want = orc.integrate_reference(before, acc[:, :3], 0.01)
    assert np.array_equal(after.view(np.uint8), want.view(np.uint8))


def test_c2_full_size(gcs, orc):
    stars = gcs.collision(131072, 1002)
    n = len(stars)
    eps2 = float(np.float32(0.02) ** 2)
    p = gcs.params(G=G, eps2=eps2, dt=0.005, theta=0.5, leaf_size=16)
    with gcs.Engine(n) as e:
        e.load(stars)
        say("C2 compute")
        e.compute_accels(p)
        cpu = check_pipeline(gcs, orc, e, stars, p)
        acc = e.accels()
        s = cpu["sorted"]
        tgt = np.linspace(0, n - 1, 4096).astype(np.int64)
        ref = orc.direct_sum_f64(s["x"], s["y"], s["z"], s["mass"], G, eps2, tgt)
        err = rel_rms(acc[tgt, :3], ref[:, :3])
        say("C2 rel rms vs FP64 direct sum: %.3g" % err)
        assert err < 1e-3
        e.step(p, 5)
        say("C2 5 steps ok")


@pytest.mark.parametrize("per,eps,dt,seed,checks", [
    (8192, 0.05, 0.01, 1001, (0, 50, 100)),              # C1, 100 steps
])
def test_pipeline_over_many_steps(gcs, orc, per, eps, dt, seed, checks):
    """C1 stepped at the bench's settings (C2's every-step check is in
    test_gpu_configs.py).  At each checkpoint the
    current state is read back and the oracle restates the frame from it:
    bounds, Morton keys, the sort's permutation, the topology and the node
    aggregates are bit-exact.  At the first, middle and last checkpoints the
    accelerations are within 1e-3 relative RMS of the FP64 direct sum (4,096
    targets)."""
    stars = gcs.collision(per, seed)
    n = len(stars)
    eps2 = float(np.float32(eps) ** 2)
    p = gcs.params(G=G, eps2=eps2, dt=dt, theta=0.5, leaf_size=4, group_size=96, team=4,
                   block_size=256, waves_per_simd=4)
    tgt = np.linspace(0, n - 1, 4096).astype(np.int64)
    direct = (checks[0], checks[len(checks) // 2], checks[-1])
    with gcs.Engine(n) as e:
        e.load(stars)
        done = 0
        for at in checks:
            if at > done:
                e.step(p, at - done)
                done = at
            cur = e.stars()
            e.compute_accels(p)
            cpu = check_pipeline(gcs, orc, e, cur, p)
            say("n=%d step %d: keys, permutation, topology and aggregates bit-exact" % (n, at))
            if at in direct:
                acc = e.accels()
                s = cpu["sorted"]
                ref = orc.direct_sum_f64(s["x"], s["y"], s["z"], s["mass"], G, eps2, tgt)
                err = rel_rms(acc[tgt, :3], ref[:, :3])
                say("n=%d step %d: rel rms vs FP64 direct sum %.3g" % (n, at, err))
                assert err < 1e-3


@pytest.mark.parametrize("half,levels", [(450000, 1), (600000, 1), (600000, 2)])
def test_large_instance_aggregates(gcs, orc, half, levels):
    """Large instances: 2 x 450,000 particles has between 1024 and 2048 nodes
    above the aggregation subtrees (the LDS top pass with two list slots per
    thread); 2 x 600,000 has more than 2048 (the global top pass, arrival
    marks in an LDS hash since the instance exceeds 2^20 particles) -- in one
    level, or in two.  Topology and aggregates bit-exact against the oracle."""
    stars = gcs.collision(half, 1006)
    n = len(stars)
    eps2 = float(np.float32(0.02) ** 2)
    p = gcs.params(G=G, eps2=eps2, dt=0.005, theta=0.5, leaf_size=8, instance_pass=levels)
    with gcs.Engine(n) as e:
        e.load(stars)
        say("large instance compute n=%d" % n)
        e.compute_accels(p)
        check_pipeline(gcs, orc, e, stars, p)
        e.step(p, 2)


def mid_level_nodes(cpu, atile=1024, mid=32768):
    """Per mid-level subtree (a subtree of more than `atile` and fewer than
    `mid` particles whose parent has at least `mid`), the number of its nodes
    of more than `atile` particles, from the oracle's tree."""
    first, last = cpu["first"], cpu["last"]
    size = last - first + 1
    parent = np.full(2 * len(first) + 1, -1)
    parent[cpu["left"]] = np.arange(len(first))
    parent[cpu["right"]] = np.arange(len(first))
    top = np.where(size > atile)[0]
    roots = [v for v in top if size[v] < mid and parent[v] >= 0 and size[parent[v]] >= mid]
    return [int(np.sum((first[top] >= first[r]) & (last[top] <= last[r]))) for r in roots]


@pytest.mark.parametrize("levels", [1, 2])
def test_forest_mid_level_subtrees(gcs, orc, levels):
    """Three collisions of 38,233-58,347 particles in one forest, whose
    mid-level subtrees hold up to 78 nodes of more than 1024 particles: each
    instance's keys, topology and aggregates bit-exact against the oracle, in
    one level and in two."""
    parts = [gcs.generate(gcs.random_spec(1004, k, 4096, 65536)) for k in (8, 24, 132)]
    counts = [len(a) for a in parts]
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=4, group_size=96,
                   instance_pass=levels)
    with gcs.Engine(sum(counts), len(parts)) as e:
        e.load(np.concatenate(parts), counts)
        say("forest of %s particles, instance pass in %d level(s)" % (counts, levels))
        e.compute_accels(p)
        raw, srt, _ = e.keys()
        trees = [e.tree(i) for i in range(len(parts))]
    most = 0
    off = 0
    for i, a in enumerate(parts):
        sl = slice(off, off + len(a))
        off += len(a)
        cpu = orc.build_tree_pipeline(a, 0.5)
        assert np.array_equal(raw[sl], cpu["keys"]) and np.array_equal(srt[sl], cpu["skeys"])
        for k in ("left", "right", "parent"):
            assert np.array_equal(trees[i][k], cpu[k]), (i, k)
        for k in ("com4", "width", "bmin", "bmax"):
            assert np.array_equal(trees[i][k].view(np.uint32), cpu["agg"][k].view(np.uint32)), \
                (i, k)
        most = max([most] + mid_level_nodes(cpu))
    say("most nodes of more than 1024 particles in a mid-level subtree: %d" % most)
    assert most > 64


def check_instances(gcs, orc, e, parts, theta=0.5):
    """Each instance of a forest (its stars in the engine's current order):
    Morton keys, sorted keys, permutation, topology and aggregates equal the
    oracle's for that instance alone."""
    raw, srt, perm = e.keys()
    off = 0
    for i, a in enumerate(parts):
        sl = slice(off, off + len(a))
        off += len(a)
        cpu = orc.build_tree_pipeline(a, theta)
        assert np.array_equal(raw[sl], cpu["keys"]), i
        assert np.array_equal(srt[sl], cpu["skeys"]) and np.array_equal(perm[sl], cpu["perm"]), i
        if len(a) == 1:
            continue
        t = e.tree(i)
        for k in ("left", "right", "parent"):
            assert np.array_equal(t[k], cpu[k]), (i, k)
        for k in ("com4", "width", "bmin", "bmax"):
            assert np.array_equal(t[k].view(np.uint32), cpu["agg"][k].view(np.uint32)), (i, k)


@pytest.mark.parametrize("case", ["cloud", "coincident", "forest"])
def test_bucket_sort_unordered_keys(gcs, orc, case):
    """The bucket sort forced on keys in no particular order (sort=2): the
    splitters fit them badly, and buckets of more than 4,096 keys are sorted
    in LDS chunks merged in global memory.  A cloud of 60,000; 20,000
    coincident particles among 10,000 others (one bucket holds every equal
    key); a forest with instances at the bucket edges (1-3, 1,535-1,537,
    3,073 particles).  Every instance's keys, permutation, topology and
    aggregates equal the oracle's."""
    if case == "cloud":
        parts = [cloud(gcs, 60000, 9)]
    elif case == "coincident":
        a = cloud(gcs, 30000, 10)
        for k, v in (("x", 0.3), ("y", -0.2), ("z", 0.1)):
            a[k][5000:25000] = v
        parts = [a]
    else:
        parts = [cloud(gcs, n, 70 + n) for n in (1, 2, 3, 1535, 1536, 1537, 3073, 1025, 20000)]
    counts = [len(a) for a in parts]
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8, sort=2)
    with gcs.Engine(sum(counts), len(parts)) as e:
        e.load(np.concatenate(parts), counts)
        say("bucket sort on unordered keys: %s, %s particles" % (case, counts))
        e.compute_accels(p)
        check_instances(gcs, orc, e, parts)


def test_bucket_sort_after_steps_and_joins(gcs, orc):
    """Automatic sort choice with step graphs: the first step after the load
    sorts by radix and stores the particles in sorted order, the graphs then
    sort every instance with buckets; an instance joined later is sorted by
    radix while the others keep buckets, and then joins them.  Checked against
    the oracle on the current state after the steps and after the join, and
    against the radix sort (sort=1) on the same state."""
    parts = [gcs.collision(n, 80 + n) for n in (768, 1536, 12000)]
    parts.append(cloud(gcs, 3073, 5))
    # A dense core among 1,000 others: after the steps its keys are still
    # nearly sorted, with buckets far narrower than the key space.
    core = cloud(gcs, 25000, 6, spread=1e-3)
    core[:1000] = cloud(gcs, 1000, 7)
    parts.append(core)
    counts = [len(a) for a in parts]
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8, use_graph=1)
    with gcs.Engine(sum(counts) + 20000, len(parts) + 1) as e:
        e.load(np.concatenate(parts), counts)
        say("bucket sort after steps: %s particles" % counts)
        e.step(p, 3)
        cur = e.stars()
        e.compute_accels(p)
        check_instances(gcs, orc, e, np.split(cur, np.cumsum(counts)[:-1]))
        acc = e.accels()
        e.compute_accels(gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8, sort=1))
        assert np.array_equal(acc.view(np.uint32), e.accels().view(np.uint32))
        joined = gcs.collision(10000, 90)
        e.add(joined)
        counts.append(len(joined))
        for steps in (0, 2):
            e.step(p, steps)
            cur = e.stars()
            e.compute_accels(p)
            say("%d steps after a join: %s particles" % (steps, counts))
            check_instances(gcs, orc, e, np.split(cur, np.cumsum(counts)[:-1]))


# ---- forests, graphs, composition ----------------------------------------------------

@pytest.mark.parametrize("team,kappa", [(1, 0.0), (2, 0.0), (4, 0.0), (8, 0.0), (1, 6.0),
                                        (4, 6.0)])
def test_forest_equals_standalone(gcs, orc, team, kappa):
    specs = [gcs.random_spec(11, k, 600, 9000) for k in range(6)]
    parts = [gcs.generate(s) for s in specs]
    parts.append(cloud(gcs, 1, 3))      # single particle instance
    parts.append(cloud(gcs, 1025, 4))   # crosses a tile boundary
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8, team=team,
                   local_kappa=kappa, block_size=64 * team if team > 1 else 256)
    counts = [len(a) for a in parts]
    with gcs.Engine(sum(counts), len(parts)) as e:
        e.load(np.concatenate(parts), counts)
        say("forest compute k=%d" % len(parts))
        e.compute_accels(p)
        acc_all = e.accels()
        raw, srt, perm = e.keys()
        trees = [e.tree(i) for i in range(len(parts))]
    off = 0
    for i, a in enumerate(parts):
        with gcs.Engine(len(a)) as e1:
            e1.load(a)
            e1.compute_accels(p)
            acc1 = e1.accels()
            r1, s1, p1 = e1.keys()
            t1 = e1.tree(0)
        sl = slice(off, off + len(a))
        assert np.array_equal(acc_all[sl].view(np.uint32), acc1.view(np.uint32)), i
        assert np.array_equal(raw[sl], r1) and np.array_equal(srt[sl], s1)
        assert np.array_equal(perm[sl], p1)
        for k in ("left", "right", "parent", "com4"):
            assert np.array_equal(trees[i][k], t1[k]), (i, k)
        off += len(a)


@pytest.mark.parametrize("team,gsize", [(1, 96), (2, 64)])
def test_ragged_forest_matches_oracle(gcs, orc, team, gsize):
    """Instances at the size edges -- one particle, two or three (the root's
    children leaves), a group's size and one more, an aggregation subtree's
    size and one either side -- walked and evaluated in one forest: each
    instance's keys and tree equal the oracle's, the forest's interaction
    counts equal the sum of the oracle's per instance, and every acceleration
    equals the oracle's FP64 sum of the same lists within 3e-5 of its terms."""
    sizes = [1, 2, 3, 5, gsize, gsize + 1, 1023, 1024, 1025, 3000]
    parts = [cloud(gcs, n, 50 + n) for n in sizes]
    eps2 = 0.0025
    p = gcs.params(G=G, eps2=eps2, dt=0.01, theta=0.5, leaf_size=4, group_size=gsize,
                   team=team, local_kappa=4.0, block_size=64 * team if team > 1 else 256)
    counts = [len(a) for a in parts]
    with gcs.Engine(sum(counts), len(parts)) as e:
        e.load(np.concatenate(parts), counts)
        say("ragged forest, team %d, groups of %d" % (team, gsize))
        e.compute_accels(p)
        acc = e.accels()
        raw, srt, perm = e.keys()
        got_counts = e.interaction_counts()
        trees = [e.tree(i) for i in range(len(parts))]
    want = np.zeros(5, np.int64)
    off = 0
    for i, a in enumerate(parts):
        n = len(a)
        sl = slice(off, off + n)
        off += n
        if n == 1:
            assert np.all(acc[sl, :3] == 0.0)
            continue
        cpu = orc.build_tree_pipeline(a, 0.5)
        assert np.array_equal(raw[sl], cpu["keys"]) and np.array_equal(srt[sl], cpu["skeys"])
        for k in ("left", "right", "parent"):
            assert np.array_equal(trees[i][k], cpu[k]), (n, k)
        s = cpu["sorted"]
        grp = orc.tree_groups(cpu["left"], cpu["right"], cpu["first"], cpu["last"], gsize)
        ref, _, cnt, aabs = orc.traverse_group(cpu["left"], cpu["right"], cpu["first"],
                                               cpu["last"], s["x"], s["y"], s["z"], s["mass"],
                                               cpu["agg"], G, eps2, gsize, 4, want_abs=True,
                                               groups=grp, local_kappa=4.0)
        want += cnt
        dev = np.abs(acc[sl, :3] - ref[:, :3]).max(1) / np.maximum(aabs, 1e-30)
        assert dev.max() < 3e-5, (n, dev.max())
    assert got_counts == [int(c) for c in want]


def test_graph_replays_match_eager(gcs):
    stars = gcs.collision(4000, 21)
    pg = gcs.params(G=G, eps2=0.0025, dt=0.01, use_graph=1, diagnostics=1)
    pe = gcs.params(G=G, eps2=0.0025, dt=0.01, use_graph=0, diagnostics=1)
    with gcs.Engine(len(stars)) as a, gcs.Engine(len(stars)) as b:
        a.load(stars)
        b.load(stars)
        for rep in range(4):
            say("graph replay %d" % rep)
            a.step(pg, 3)
            b.step(pe, 3)
            _ = b.stars()  # eager work between replays
            assert np.array_equal(a.stars().view(np.uint8), b.stars().view(np.uint8)), rep
            assert np.array_equal(a.diagnostics(), b.diagnostics())


def test_instance_remove_and_add(gcs):
    parts = [gcs.collision(n, 30 + n) for n in (700, 1500, 900)]
    p = gcs.params(G=G, eps2=0.0025, dt=0.01)
    with gcs.Engine(20000, 8) as e:
        e.load(np.concatenate(parts), [len(a) for a in parts])
        e.step(p, 2)
        keep0, keep2 = e.instance(0), e.instance(2)
        e.remove(1)
        assert e.count() == (2, len(parts[0]) + len(parts[2]))
        assert np.array_equal(e.instance(0), keep0) and np.array_equal(e.instance(2), keep2)
        new = gcs.collision(1200, 99)
        iid = e.add(new)
        assert iid == 3
        say("composition step")
        e.step(p, 3)
        got = [e.instance(i) for i in (0, 2, 3)]
    for start, g in zip((keep0, keep2, new), got):
        with gcs.Engine(len(start)) as e1:
            e1.load(start)
            e1.step(p, 3)
            assert np.array_equal(e1.stars().view(np.uint8), g.view(np.uint8))


# ---- precision and diagnostics --------------------------------------------------------

def test_compensated_positions(gcs):
    stars = gcs.collision(3000, 41)
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, precision=gcs.PREC_COMPENSATED)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        say("compensated steps")
        e.step(p, 20)
        s = e.stars()
        lo = e.position_lo()
    hi = np.stack([s["x"], s["y"], s["z"]], 1).astype(np.float64)
    ulp = np.spacing(np.abs(hi).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(lo) <= 0.5 * ulp + 1e-30)
    assert np.any(lo != 0)


def test_diagnostics_match_host_sums(gcs, orc):
    stars = gcs.collision(2000, 42)
    eps2 = 0.0025
    p = gcs.params(G=G, eps2=eps2, dt=0.01, diagnostics=1, use_graph=0)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        e.compute_accels(p)
        s = e.stars()
        acc = e.accels()
        e.step(p, 1)
        d = e.diagnostics()
    m = s["mass"].astype(np.float64)
    v = np.stack([s["vx"], s["vy"], s["vz"]], 1).astype(np.float64) + \
        0.5 * 0.01 * acc[:, :3].astype(np.float64)
    ke = 0.5 * (m * (v ** 2).sum(1)).sum()
    ref = orc.direct_sum_f64(s["x"], s["y"], s["z"], s["mass"], G, eps2)
    pe = 0.5 * (m * ref[:, 3]).sum()
    assert np.isclose(d[0], ke, rtol=1e-9)
    assert np.isclose(d[1], pe, rtol=1e-3)
    assert np.allclose(d[2:], (m[:, None] * v).sum(0), rtol=1e-9, atol=1e-12)


# ---- edge cases and errors ---------------------------------------------------------

def test_coincident_particles(gcs, orc):
    # All at one point: every key equal, the tree splits by index only, and
    # every pair is at zero separation (zero force).
    st = np.zeros(300, gcs.STAR)
    st["x"], st["y"], st["z"], st["mass"] = 1.5, -2.0, 0.25, 1.0
    p = gcs.params(G=G, eps2=0.01, dt=0.01)
    with gcs.Engine(300) as e:
        e.load(st)
        e.compute_accels(p)
        raw, srt, perm = e.keys()
        assert np.all(raw == raw[0]) and np.array_equal(perm, np.arange(300))
        assert raw[0] == orc.morton(1.5, -2.0, 0.25, -2.0, 1.5)
        t = e.tree(0)
        el, er, ep = orc.karras_build(srt)
        assert np.array_equal(t["left"], el) and np.array_equal(t["parent"], ep)
        assert np.all(e.accels()[:, :3] == 0)


def test_stack_overflow_is_reported(gcs):
    stars = cloud(gcs, 2000, 9)
    p = gcs.reference_params(G, 1e-6, 0.01, -6.0, 6.0, stack_depth=2)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        with pytest.raises(gcs.GcsError) as ex:
            e.step(p, 1)
        assert ex.value.code == -6
        e.step(gcs.params(G=G, eps2=0.01, dt=0.01), 1)  # the engine still works


def test_autotune_picks_a_candidate_and_keeps_the_state(gcs):
    gcs.autotune_cache_clear()
    stars = gcs.collision(8192, 44)
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, use_graph=1)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        say("autotune group")
        best, ms = e.autotune(p, reps=2)
        assert ms > 0
        assert best.leaf_size in (4, 8, 12, 16) and best.block_size in (64, 128, 256)
        assert best.waves_per_simd in (2, 3, 4) and best.use_graph == 1
        assert best.group_size in (48, 64, 96, 128)
        assert best.theta == p.theta and best.eps2 == p.eps2
        # Kept per (kind, theta, form, precision, N bucket, k bucket): the
        # same answer again without timing.
        again, ms2 = e.autotune(p, reps=2)
        assert (again.group_size, again.leaf_size, again.block_size, again.team,
                again.waves_per_simd, ms2) == (best.group_size, best.leaf_size,
                                               best.block_size, best.team,
                                               best.waves_per_simd, ms)
        # positions unchanged (re-sorted only)
        after = np.sort(e.stars(), order=["x", "y", "z"])
        before = np.sort(stars, order=["x", "y", "z"])
        assert np.array_equal(after.view(np.uint8), before.view(np.uint8))
        e.step(best, 2)
        say("autotune reference")
        r = gcs.reference_params(G, 0.0025, 0.01, -120.0, 120.0)
        best_r, ms_r = e.autotune(r, reps=1)
        assert best_r.stack_depth in (64, 128, 256, 1024) and ms_r > 0
        e.step(best_r, 1)


def test_checkpoint_resume_is_bit_exact(gcs):
    a = gcs.generate(gcs.random_spec(21, 0, 5000, 5000))
    b = gcs.generate(gcs.random_spec(21, 1, 3000, 3000))
    p = gcs.params(G=G, eps2=4e-4, dt=0.005, precision=gcs.PREC_COMPENSATED, use_graph=1)
    with gcs.Engine(len(a) + len(b), 2) as e:
        e.load(np.concatenate([a, b]), [len(a), len(b)])
        say("checkpoint: 7 steps, checkpoint, 9 steps")
        e.step(p, 7)
        ck = e.checkpoint(1)
        assert len(ck) == gcs.lib().gcs_checkpoint_bytes(len(b))
        e.step(p, 9)
        want = e.instance(1)
        n1, off1 = e.info(1)
        want_lo = e.position_lo()[off1:off1 + n1]
    with gcs.Engine(len(b)) as e2:
        say("resume: 9 steps")
        iid = e2.restore(ck)
        e2.step(p, 9)
        got = e2.instance(iid)
        got_lo = e2.position_lo()
    assert np.array_equal(want.view(np.uint8), got.view(np.uint8))
    assert np.array_equal(want_lo.view(np.uint32), got_lo.view(np.uint32))
    assert np.any(got_lo != 0)


def test_max_accel_and_adaptive_step(gcs):
    stars = gcs.collision(4000, 45)
    eps2 = float(np.float32(0.02) ** 2)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        fixed = gcs.params(G=G, eps2=eps2, dt=0.005, use_graph=0)
        e.compute_accels(fixed)
        acc = e.accels()
        amax = float(np.sqrt((acc[:, :3].astype(np.float32) ** 2).sum(1)).max())
        say("fixed step: max |a| and simulated time")
        e.step(fixed, 1)
        assert abs(e.max_accel() - amax) <= 1e-6 * amax
        assert e.sim_time() == np.float64(np.float32(0.005))
        # Adaptive: dt = min(dt_max, sqrt(2 eta eps / max|a|)) on that step's
        # accelerations (positions of the first step's end).
        e.compute_accels(fixed)
        a2 = e.accels()
        m2 = np.float32(np.sqrt((a2[:, :3].astype(np.float32) ** 2).sum(1)).max())
        # eta chosen so the adaptive step is about half the maximum
        eta = float(0.25 * 0.005 ** 2 * float(m2) / (2 * np.sqrt(eps2)))
        want = min(np.float32(0.005), np.float32(np.sqrt(np.float32(2 * eta) *
                                                         np.float32(np.sqrt(eps2)) / m2)))
        say("adaptive step")
        e.step(gcs.params(G=G, eps2=eps2, dt=0.005, dt_eta=eta, use_graph=0), 1)
        got = e.sim_time() - np.float64(np.float32(0.005))
        assert want < np.float32(0.005)
        assert abs(got - float(want)) <= 1e-6 * float(want)
        t0 = e.sim_time()
        e.step(gcs.params(G=G, eps2=eps2, dt=0.005, dt_eta=eta, use_graph=1), 3)
        assert e.sim_time() > t0


@pytest.mark.parametrize("mode", ["pipeline", "via-shim"])
def test_bake_driver_writes_reference_frames(gcs, tmp_path, mode):
    """gcs_bake (the driver) writes simulation.bin: N x 16-byte PlaybackStar
    frames in device order.  Its default mode runs the MI355X pipeline
    (gcs_bake: device bounds, device pack, async 16-byte copies); --via-shim
    runs the reference's frame through the cuda_engine.h functions (host
    bounds, host sort, cuda_physics_step, host pack).  Both must equal the
    frames of the engine stepped with the reference's semantics and host
    bounds, read back after each step."""
    exe = os.path.join(os.path.dirname(gcs.LIB_PATH), "gcs_bake")
    out = tmp_path / "simulation.bin"
    n_per = 3000
    say("bake driver, %s" % mode)
    r = subprocess.run([exe, "5", "--n", str(n_per), "--seed", "5", "--eps2", "0.0025",
                        "--out", str(out)] + (["--via-shim"] if mode == "via-shim" else []),
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    assert '"frames": 5' in r.stdout
    data = np.fromfile(out, gcs.PLAYBACK_STAR)
    assert len(data) == 5 * 2 * n_per
    stars = gcs.collision(n_per, 5)
    with gcs.Engine(len(stars)) as e:
        e.load(stars)
        host = stars
        for f in range(5):
            mn = min(host["x"].min(), host["y"].min(), host["z"].min())
            mx = max(host["x"].max(), host["y"].max(), host["z"].max())
            e.step(gcs.reference_params(G, 0.0025, 0.01, float(mn), float(mx)), 1)
            host = e.stars()
            pb = e.playback()
            assert np.array_equal(pb.view(np.uint8), data[f * 2 * n_per:(f + 1) * 2 * n_per]
                                  .view(np.uint8)), f


def test_bake_entry_point_frames(gcs, tmp_path):
    """gcs_bake on a forest of two collisions with the group traversal: each
    frame equals gcs_read_playback after the same step of an engine stepped
    one frame at a time (frames of 7 and 12 steps cover both buffers and
    their reuse)."""
    parts = [gcs.collision(2500, 71), gcs.collision(1200, 72)]
    counts = [len(a) for a in parts]
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8, use_graph=1,
                   precision=gcs.PREC_COMPENSATED)
    for frames in (7, 12):
        out = tmp_path / ("bake%d.bin" % frames)
        with gcs.Engine(sum(counts), 2) as e:
            e.load(np.concatenate(parts), counts)
            say("gcs_bake %d frames" % frames)
            sec = e.bake(p, frames, out)
            assert sec > 0
        data = np.fromfile(out, gcs.PLAYBACK_STAR)
        n = sum(counts)
        assert len(data) == frames * n
        with gcs.Engine(n, 2) as e:
            e.load(np.concatenate(parts), counts)
            for f in range(frames):
                e.step(p, 1)
                assert np.array_equal(e.playback().view(np.uint8),
                                      data[f * n:(f + 1) * n].view(np.uint8)), (frames, f)


# ---- reference traversal thresholds, root record, compensated drift -----------------

@pytest.mark.parametrize("case", ["open-0.99", "open-1.01", "skip-0.98", "skip-1.02"])
def test_reference_thresholds_on_gpu(gcs, orc, case):
    """The threshold cases that pin the oracle's reference traversal
    (tests/test_oracle.py: a node taken iff width^2 < d^2 + eps^2, a leaf
    skipped iff d^2 + eps^2 <= 1.01 eps^2), run through the HIP reference
    traversal: the same interactions and accelerations."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import test_oracle as to
    kind, over = case.split("-")
    if kind == "open":
        st, _, (mn, mx), eps2 = to.reference_opening_case(float(over))
    else:
        st, (mn, mx), eps2 = to.reference_skip_case(float(over))
    stars = np.zeros(3, gcs.STAR)
    for k in ("x", "y", "z", "mass"):
        stars[k] = st[k]
    p = gcs.reference_params(G, eps2, 0.01, mn, mx, use_graph=0)
    with gcs.Engine(3) as e:
        e.load(stars)
        say("reference threshold case %s" % case)
        e.compute_accels(p)
        cpu = check_pipeline(gcs, orc, e, stars, p, theta=1.0, bounds=(mn, mx))
        acc = e.accels()
        s = cpu["sorted"]
        want, _, cnt = orc.traverse_reference(cpu["left"], cpu["right"], s["x"], s["y"], s["z"],
                                              cpu["agg"]["com4"], cpu["agg"]["width"], G, eps2)
        assert e.interactions()[1] == cnt
        assert np.allclose(acc[:, :3], want[:, :3], rtol=1e-5, atol=0)


def test_root_record_matches_oracle(gcs, orc):
    """gcs_read_root copies the root's (COM, mass) -- bit-exact against the
    oracle's root aggregate for each instance of a forest, a one-particle
    instance included (its leaf, the position as built), before and after a
    step; and cuda_get_center_of_mass returns it through the reference's
    C++ entry points (cuda_engine.cu:365-376)."""
    parts = [gcs.collision(1500, 61), cloud(gcs, 1, 62), cloud(gcs, 700, 63)]
    counts = [len(a) for a in parts]
    p = gcs.params(G=G, eps2=0.0025, dt=0.01, theta=0.5, leaf_size=8)
    with gcs.Engine(sum(counts), 3) as e:
        e.load(np.concatenate(parts), counts)
        say("root records of a forest")
        e.compute_accels(p)
        roots = [e.root(i) for i in range(3)]
        e.step(p, 1)  # re-sorts unchanged keys: the same tree, then the drift
        after = [e.root(i) for i in range(3)]
    for i, a in enumerate(parts):
        cpu = orc.build_tree_pipeline(a, 0.5)
        want = cpu["agg"]["com4"][0]
        assert np.array_equal(roots[i].view(np.uint32), want.view(np.uint32)), i
        assert np.array_equal(after[i].view(np.uint32), want.view(np.uint32)), i
    # Through the cuda_engine.h shim (one process-wide engine).
    import ctypes
    L = gcs.lib()
    init = L._Z9cuda_initi
    init.argtypes, init.restype = [ctypes.c_int], ctypes.c_bool
    up = L._Z25cuda_upload_initial_starsPK4Stari
    up.argtypes, up.restype = [ctypes.c_void_p, ctypes.c_int], None
    stepf = L._Z17cuda_physics_stepP4Starifffffb
    stepf.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_float] * 5 + [ctypes.c_bool]
    stepf.restype = ctypes.c_float
    com = L._Z23cuda_get_center_of_massPfS_S_
    com.argtypes, com.restype = [ctypes.c_void_p] * 3, None
    cleanup = L._Z12cuda_cleanupv
    cleanup.argtypes, cleanup.restype = [], None
    stars = parts[0]
    n = len(stars)
    mn = float(min(stars["x"].min(), stars["y"].min(), stars["z"].min()))
    mx = float(max(stars["x"].max(), stars["y"].max(), stars["z"].max()))
    out = np.zeros(n, gcs.STAR)
    c = [ctypes.c_float(1.0) for _ in range(3)]
    assert init(n)
    try:
        com(*(ctypes.byref(v) for v in c))
        assert [v.value for v in c] == [0.0, 0.0, 0.0]  # nothing built yet
        up(stars.ctypes.data, n)
        say("cuda_physics_step through the shim")
        dt = stepf(out.ctypes.data, n, 1.0, 0.0025, 0.01, mn, mx, False)
        assert dt == np.float32(0.01)
        com(*(ctypes.byref(v) for v in c))
    finally:
        cleanup()
    cpu = orc.build_tree_pipeline(stars, 1.0, mn, mx)
    want = cpu["agg"]["com4"][0]
    got = np.array([v.value for v in c], np.float32)
    assert np.array_equal(got.view(np.uint32), want[:3].view(np.uint32))


def test_compensated_drift_closed_form(gcs):
    """Free particles (one per instance: no force acts) drift at constant
    velocity, so after k steps the exact position is x0 + k (v dt).  With
    compensated positions hi + lo tracks it in FP64 to ~1e-12 relative over
    1,000 steps (each step rounds the FP32 low part, ~2^-24 of an FP32 ulp);
    plain FP32 positions drift by ~1e-6 relative.  A kernel that drops or
    resets the low part fails the first bound."""
    k, steps, dt = 96, 1000, np.float32(0.005)
    r = np.random.default_rng(46)
    st = np.zeros(k, gcs.STAR)
    for f in ("x", "y", "z"):
        st[f] = r.uniform(40.0, 160.0, k).astype(np.float32) * r.choice([-1, 1], k)
    for f in ("vx", "vy", "vz"):
        st[f] = r.uniform(-1.5, 1.5, k).astype(np.float32)
    st["mass"] = 1.0
    x0 = np.stack([st["x"], st["y"], st["z"]], 1).astype(np.float64)
    v = np.stack([st["vx"], st["vy"], st["vz"]], 1).astype(np.float64)
    exact = x0 + steps * (v * np.float64(dt))
    err = {}
    for prec in (gcs.PREC_PLAIN, gcs.PREC_COMPENSATED):
        p = gcs.params(G=G, eps2=0.0025, dt=float(dt), precision=prec, use_graph=1)
        with gcs.Engine(k, k) as e:
            e.load(st, [1] * k)
            say("free drift, precision %d, %d steps" % (prec, steps))
            e.step(p, steps)
            s = e.stars()
            lo = e.position_lo() if prec == gcs.PREC_COMPENSATED else 0.0
        pos = np.stack([s["x"], s["y"], s["z"]], 1).astype(np.float64) + lo
        assert np.array_equal(s["vx"], st["vx"])  # no force: velocities unchanged
        err[prec] = np.abs(pos - exact).max() / np.abs(exact).max()
    say("free drift max relative error: plain %.3g, compensated %.3g"
        % (err[gcs.PREC_PLAIN], err[gcs.PREC_COMPENSATED]))
    assert err[gcs.PREC_COMPENSATED] < 1e-11
    assert err[gcs.PREC_PLAIN] > 1e-8
