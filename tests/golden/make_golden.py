"""Writes the golden fixtures in this folder from the reference's own code.

The reference ships no tests or fixtures, so the vectors here come from its
sources compiled in this container by oracle/Makefile into oracle/_ref/
(morton.h get_morton_code; engine.cpp + gravity_aggregator.cpp driven in the
order of the reference's CPU frame, main.cpp:152-253).  Run from the repo root
after `make -f oracle/Makefile all`:

    python3 tests/golden/make_golden.py

morton_ref.npz   cases of (x, y, z, min, max) -> 63-bit keys
octree_ref.npz   one reference CPU frame on a 2,048-particle input: the
                 accelerations (Morton-sorted order) and the state after the
                 kick-drift, G = 1, eps^2 = 0.0025, dt = 0.01
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402


def morton_cases():
    r = np.random.default_rng(20240601)
    cases = []
    for mn, mx in ((-7.25, 3.5), (-60.0, 60.0), (0.0, 1e-3), (-1e6, 2.5e6)):
        n = 1024
        x, y, z = (r.uniform(mn, mx, n).astype(np.float32) for _ in range(3))
        # points on and next to the bounds, and on the cell grid
        f32 = np.float32
        edge = np.array([mn, mx, np.nextafter(f32(mx), f32(-np.inf)),
                         np.nextafter(f32(mn), f32(np.inf)), (mn + mx) / 2], np.float32)
        x[:5], y[5:10], z[10:15] = edge, edge, edge
        x = np.clip(x, f32(mn), f32(mx))
        y = np.clip(y, f32(mn), f32(mx))
        z = np.clip(z, f32(mn), f32(mx))
        cases.append((x, y, z, mn, mx, oracle.ref_morton(x, y, z, mn, mx)))
    return cases


def octree_input(n=2048, seed=7):
    r = np.random.default_rng(seed)
    st = np.zeros(n, oracle.STAR)
    h = n // 2
    for k, c in enumerate(((-3.0, 0.5, 0.0), (4.0, -1.0, 0.5))):
        sl = slice(k * h, (k + 1) * h)
        rad = 1.5 * r.standard_normal((h, 3)) / np.sqrt(1 + r.uniform(0, 4, (h, 1)))
        for a, key in enumerate("xyz"):
            st[key][sl] = (c[a] + rad[:, a]).astype(np.float32)
        st["vx"][sl] = (0.1 - 0.2 * k) + 0.05 * r.standard_normal(h)
        st["vy"][sl] = 0.05 * r.standard_normal(h)
        st["vz"][sl] = 0.05 * r.standard_normal(h)
    st["mass"] = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    st["type"] = r.integers(0, 3, n)
    st["r"], st["g"], st["b"] = 10, 20, 30
    return st


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref not built: run make -f oracle/Makefile all")
    cs = morton_cases()
    out = {}
    for i, (x, y, z, mn, mx, k) in enumerate(cs):
        out["x%d" % i], out["y%d" % i], out["z%d" % i] = x, y, z
        out["bounds%d" % i] = np.array([mn, mx], np.float32)
        out["keys%d" % i] = k
    np.savez_compressed(os.path.join(HERE, "morton_ref.npz"), cases=len(cs), **out)

    st = octree_input()
    G, eps2, dt = 1.0, 0.0025, 0.01
    after, acc, _ = oracle.ref_octree_step(st, G, eps2, dt, integrate=True, want_acc=True)
    np.savez_compressed(os.path.join(HERE, "octree_ref.npz"), stars=st, acc=acc, after=after,
                        params=np.array([G, eps2, dt], np.float32))
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
