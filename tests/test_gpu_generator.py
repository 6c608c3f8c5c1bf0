"""Initial conditions generated on the device (gcs_load_generated) against the
host generator (gcs_generate), which the CPU tests pin statistically.

Both run the same sampling code on the same counter-based streams, in FP64
with no multiply-add contraction.  Bar: types and colours identical; float
fields identical except where the device's log/cos/sin/pow differ from the
host's in the last bit -- at most 1e-4 of the values, each within 4 float
ulps -- and the engine steps the generated state like a loaded one.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FLOATS = ("x", "y", "z", "vx", "vy", "vz", "mass")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare(dev, host, what):
    assert len(dev) == len(host)
    for k in ("type", "r", "g", "b"):
        assert np.array_equal(dev[k], host[k]), (what, k)
    bad = 0
    for k in FLOATS:
        a = dev[k].astype(np.float32)
        b = host[k].astype(np.float32)
        diff = a != b
        bad += int(diff.sum())
        if diff.any():
            ulp = np.spacing(np.maximum(np.abs(a[diff]), np.abs(b[diff])))
            assert np.all(np.abs(a[diff] - b[diff]) <= 4 * ulp), (what, k)
    assert bad <= max(2, 1e-4 * len(dev) * len(FLOATS)), (what, bad)
    return bad


def host_records(gcs, specs):
    return np.concatenate([gcs.generate(s) for s in specs])


def test_default_collision_matches_host(gcs):
    import ctypes
    from gcs_amd import _lib as L
    spec = L.CollisionSpec()
    L.lib().gcs_spec_default(8192, 1001, ctypes.byref(spec))
    say("default collision 2 x 8192, seed 1001")
    with gcs.Engine(16384) as e:
        e.load_generated([spec])
        dev = e.stars()
    bad = compare(dev, host_records(gcs, [spec]), "default")
    say("values differing in the last bits:", bad)


def test_forest_of_random_collisions_matches_host(gcs):
    specs = [gcs.random_spec(5, k, 4096, 65536) for k in range(6)]
    # Edge shapes: a one-particle galaxy, an empty galaxy, odd sizes.
    import ctypes
    from gcs_amd import _lib as L
    for n_a, n_b, seed in ((1, 0, 3), (0, 7, 4), (1001, 3, 5)):
        s = L.CollisionSpec()
        L.lib().gcs_spec_default(1, seed, ctypes.byref(s))
        s.n_a, s.n_b = n_a, n_b
        specs.append(s)
    total = sum(s.n_a + s.n_b for s in specs)
    say("forest of", len(specs), "collisions,", total, "particles")
    with gcs.Engine(total, len(specs)) as e:
        e.load_generated(specs)
        dev = e.stars()
    compare(dev, host_records(gcs, specs), "forest")


def test_generated_state_steps_like_loaded(gcs):
    specs = [gcs.random_spec(9, k, 2048, 8192) for k in range(3)]
    host = host_records(gcs, specs)
    counts = [s.n_a + s.n_b for s in specs]
    p = gcs.params(eps2=float(np.float32(0.02) ** 2), dt=0.005, theta=0.5, leaf_size=8,
                   use_graph=0)
    say("3 collisions, generated vs loaded, 3 steps")
    with gcs.Engine(sum(counts), 3) as a, gcs.Engine(sum(counts), 3) as b:
        a.load_generated(specs)
        dev = a.stars()
        b.load(dev, counts)  # the device records, loaded from the host
        a.step(p, 3)
        b.step(p, 3)
        assert a.stars().tobytes() == b.stars().tobytes()
    compare(dev, host, "stepped")


def test_bad_specs_rejected(gcs):
    import ctypes
    from gcs_amd import _lib as L
    s = L.CollisionSpec()
    L.lib().gcs_spec_default(0, 1, ctypes.byref(s))
    with gcs.Engine(1024) as e:
        with pytest.raises(Exception):
            e.load_generated([s])  # no particles
        s.n_a = 2048
        with pytest.raises(Exception):
            e.load_generated([s])  # over capacity


def test_added_instance_matches_host(gcs):
    """gcs_instance_add_generated: only the new instance is generated; the
    resident ones keep their records."""
    first = [gcs.random_spec(21, k, 3000, 9000) for k in range(2)]
    extra = gcs.random_spec(22, 0, 5000, 20000)
    n0 = sum(s.n_a + s.n_b for s in first)
    say("forest of 2 + 1 added collision")
    with gcs.Engine(n0 + extra.n_a + extra.n_b, 3) as e:
        e.load_generated(first)
        before = e.stars()
        iid = e.add_generated(extra)
        assert iid == 2
        after = e.stars()
    assert after[:n0].tobytes() == before.tobytes()
    compare(after[n0:], host_records(gcs, [extra]), "added")
