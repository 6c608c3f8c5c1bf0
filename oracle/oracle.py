"""ctypes view of the CPU oracle (liboracle.so) and of the compiled reference
octree (_ref/libref_octree*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker; never by the product.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_octree.so")
REF_FAST_SO = os.path.join(HERE, "_ref", "libref_octree_fast.so")

STAR = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("vx", "<f4"), ("vy", "<f4"),
                 ("vz", "<f4"), ("mass", "<f4"), ("type", "u1"), ("r", "u1"), ("g", "u1"),
                 ("b", "u1")])

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F = ctypes.c_float
_D = ctypes.c_double

_SIG = {
    "orc_expand_bits": ([ctypes.c_uint32], ctypes.c_uint64),
    "orc_f2u_sat": ([_F], ctypes.c_uint32),
    "orc_morton": ([_F, _F, _F, _F, _F], ctypes.c_uint64),
    "orc_morton_n": ([_P, _P, _P, _I64, _F, _F, _P], None),
    "orc_bounds_cube": ([_P, _P, _P, _I64, _P, _P], None),
    "orc_stable_sort_pairs": ([_P, _I64, _P, _P], None),
    "orc_karras_build": ([_P, _I32, _P, _P, _P], None),
    "orc_radix_tree_definition": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "orc_node_ranges": ([_I32, _P, _P, _P, _P], None),
    "orc_aggregate": ([_I32, _P, _P, _P, _P, _P, _P, _F, _I32, _P, _P, _P, _P, _P, _P], None),
    "orc_traverse_reference": ([_I32, _P, _P, _P, _P, _P, _P, _P, _F, _F, _P, _P, _P], _I64),
    "orc_traverse_group": ([_I32, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _F,
                            _F, _P, _P, _P, _P, _P, _I32, _P, _F, _F], None),
    "orc_tree_groups": ([_I32, _P, _P, _P, _P, _I32, _P], _I32),
    "orc_direct_sum_f64": ([_I64, _P, _P, _P, _P, _I64, _P, _D, _D, _P], None),
    "orc_integrate_reference": ([_I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _F],
                                None),
    "orc_octree_step": ([_P, _I64, _F, _F, _F, _F, _I32, _I32, _P, _P], ctypes.c_int),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError("oracle not built: run make -f oracle/Makefile")
        L = ctypes.CDLL(ORACLE_SO)
        for k, (a, r) in _SIG.items():
            f = getattr(L, k)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def P(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def c32(a):
    return np.ascontiguousarray(a, np.float32)


def expand_bits(v):
    return int(lib().orc_expand_bits(v))


def morton(x, y, z, mn, mx):
    return int(lib().orc_morton(x, y, z, mn, mx))


def morton_n(x, y, z, mn, mx):
    x, y, z = c32(x), c32(y), c32(z)
    k = np.zeros(len(x), np.uint64)
    lib().orc_morton_n(P(x), P(y), P(z), len(x), mn, mx, P(k))
    return k


def bounds_cube(x, y, z):
    x, y, z = c32(x), c32(y), c32(z)
    a, b = ctypes.c_float(), ctypes.c_float()
    lib().orc_bounds_cube(P(x), P(y), P(z), len(x), ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def stable_sort_pairs(keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    ko = np.zeros(len(keys), np.uint64)
    po = np.zeros(len(keys), np.int32)
    lib().orc_stable_sort_pairs(P(keys), len(keys), P(ko), P(po))
    return ko, po


def _tree_arrays(n):
    m = max(2 * n - 1, 0)
    return np.zeros(m, np.int32), np.zeros(m, np.int32), np.zeros(m, np.int32)


def karras_build(skeys):
    skeys = np.ascontiguousarray(skeys, np.uint64)
    l, r, p = _tree_arrays(len(skeys))
    lib().orc_karras_build(P(skeys), len(skeys), P(l), P(r), P(p))
    return l, r, p


def radix_tree_definition(skeys):
    skeys = np.ascontiguousarray(skeys, np.uint64)
    l, r, p = _tree_arrays(len(skeys))
    rc = lib().orc_radix_tree_definition(P(skeys), len(skeys), P(l), P(r), P(p))
    assert rc == 0
    return l, r, p


def node_ranges(n, left, right):
    f = np.zeros(2 * n - 1, np.int32)
    la = np.zeros(2 * n - 1, np.int32)
    lib().orc_node_ranges(n, P(left), P(right), P(f), P(la))
    return f, la


def aggregate(left, right, x, y, z, m, theta, offset=False):
    n = len(x)
    x, y, z, m = c32(x), c32(y), c32(z), c32(m)
    k = 2 * n - 1
    com4 = np.zeros((k, 4), np.float32)
    width = np.zeros(k, np.float32)
    bmin = np.zeros((k, 3), np.float32)
    bmax = np.zeros((k, 3), np.float32)
    crit = np.zeros(k, np.float32)
    quad = np.zeros((k, 7), np.float32)
    inv_theta = np.float32(np.inf) if theta == 0 else np.float32(1.0) / np.float32(theta)
    lib().orc_aggregate(n, P(left), P(right), P(x), P(y), P(z), P(m), float(inv_theta),
                        1 if offset else 0, P(com4),
                        P(width), P(bmin), P(bmax), P(crit), P(quad))
    return dict(com4=com4, width=width, bmin=bmin, bmax=bmax, crit=crit, quad=quad)


def traverse_reference(left, right, x, y, z, com4, width, G, eps2):
    n = len(x)
    x, y, z = c32(x), c32(y), c32(z)
    acc = np.zeros((n, 4), np.float64)
    msum = np.zeros(n, np.float64)
    mskip = np.zeros(n, np.float64)
    cnt = lib().orc_traverse_reference(n, P(x), P(y), P(z), P(left), P(right),
                                       P(np.ascontiguousarray(com4, np.float32)),
                                       P(np.ascontiguousarray(width, np.float32)), G, eps2,
                                       P(acc), P(msum), P(mskip))
    return acc, (msum, mskip), int(cnt)


def traverse_group(left, right, first, last, x, y, z, m, agg, G, eps2, group=64, leaf=16,
                   quad=True, want_abs=False, gcost=None, groups=None, mono_far=0.0,
                   local_kappa=0.0):
    """counts: [node interactions evaluated per target, particle interactions,
    of the nodes those taken as monopoles (beyond mono_far times the opening
    radius; 0: none), (group, node) pairs taken through local expansions
    (nodes beyond local_kappa group radii from the group's centre; 0: none),
    the target interactions those stand for]."""
    n = len(x)
    x, y, z, m = c32(x), c32(y), c32(z), c32(m)
    acc = np.zeros((n, 4), np.float64)
    msum = np.zeros(n, np.float64)
    counts = np.zeros(5, np.int64)
    far2 = float(np.float32(mono_far) * np.float32(mono_far)) if mono_far > 0 else float("inf")
    kappa2 = (float(np.float32(local_kappa) * np.float32(local_kappa)) if local_kappa > 0
              else float("inf"))
    aabs = np.zeros(n, np.float64) if want_abs else None
    lib().orc_traverse_group(n, P(x), P(y), P(z), P(m), P(left), P(right), P(first), P(last),
                             P(agg["com4"]), P(agg["crit"]), P(agg["quad"]) if quad else None,
                             group, leaf, G, eps2, P(acc), P(counts), P(msum), P(aabs),
                             P(gcost), 0 if groups is None else len(groups),
                             None if groups is None else
                             P(np.ascontiguousarray(groups, np.int32)), far2, kappa2)
    if want_abs:
        return acc, msum, counts, aabs
    return acc, msum, counts


def tree_groups(left, right, first, last, gmax=64):
    n = (len(left) + 1) // 2
    out = np.zeros((max(n, 1), 2), np.int32)
    k = lib().orc_tree_groups(n, P(left), P(right), P(first), P(last), gmax, P(out))
    return out[:k].copy()


def direct_sum_f64(x, y, z, m, G, eps2, targets=None):
    n = len(x)
    xd, yd, zd, md = (np.ascontiguousarray(a, np.float64) for a in (x, y, z, m))
    if targets is None:
        targets = np.arange(n, dtype=np.int64)
    t = np.ascontiguousarray(targets, np.int64)
    acc = np.zeros((len(t), 4), np.float64)
    lib().orc_direct_sum_f64(n, P(xd), P(yd), P(zd), P(md), len(t), P(t), G, eps2, P(acc))
    return acc


def integrate_reference(stars, acc3, dt):
    """Kick-drift + colours on a copy of the STAR records; acc3 (n,3) float32."""
    s = np.array(stars, dtype=STAR, copy=True)
    cols = {k: np.ascontiguousarray(s[k]) for k in STAR.names}
    a = np.ascontiguousarray(acc3, np.float32)
    ax, ay, az = (np.ascontiguousarray(a[:, k]) for k in range(3))
    lib().orc_integrate_reference(len(s), P(cols["x"]), P(cols["y"]), P(cols["z"]),
                                  P(cols["vx"]), P(cols["vy"]), P(cols["vz"]), P(cols["r"]),
                                  P(cols["g"]), P(cols["b"]), P(cols["type"]), P(ax), P(ay),
                                  P(az), dt)
    for k in STAR.names:
        s[k] = cols[k]
    return s


def octree_step(stars, G, eps2, dt, theta=0.8, leaf_cap=32, integrate=True, want_acc=False):
    """One CPU frame (the restatement).  Returns (stars_after, acc or None,
    phase seconds)."""
    s = np.array(stars, dtype=STAR, copy=True)
    acc = np.zeros((len(s), 3), np.float32) if want_acc else None
    ph = np.zeros(6, np.float64)
    rc = lib().orc_octree_step(P(s), len(s), G, eps2, dt, theta, leaf_cap, 1 if integrate else 0,
                               P(acc), P(ph))
    if rc != 0:
        raise RuntimeError("octree step failed")
    return s, acc, ph


_ref = {}


def ref_available(fast=False):
    return os.path.exists(REF_FAST_SO if fast else REF_SO)


def _ref_lib(fast=False):
    path = REF_FAST_SO if fast else REF_SO
    if path not in _ref:
        L = ctypes.CDLL(path)
        L.ref_octree_step.argtypes = [_P, ctypes.c_int, _F, _F, _F, ctypes.c_int, _P, _P]
        L.ref_octree_step.restype = ctypes.c_int
        L.ref_star_size.restype = ctypes.c_int
        L.ref_morton_codes.argtypes = [_P, _P, _P, ctypes.c_int, _F, _F, _P]
        L.ref_morton_codes.restype = None
        assert L.ref_star_size() == 32
        _ref[path] = L
    return _ref[path]


def ref_morton(x, y, z, mn, mx):
    """The reference's compiled get_morton_code (morton.h:21-28) on n points."""
    x, y, z = c32(x), c32(y), c32(z)
    out = np.zeros(len(x), np.uint64)
    _ref_lib().ref_morton_codes(P(x), P(y), P(z), len(x), mn, mx, P(out))
    return out


def ref_octree_step(stars, G, eps2, dt, integrate=True, want_acc=False, fast=False):
    """One CPU frame of the reference's own octree code (theta = 0.8 and leaf
    cap 32 as compiled into gravity_aggregator.cpp / engine.cpp)."""
    L = _ref_lib(fast)
    s = np.array(stars, dtype=STAR, copy=True)
    acc = np.zeros((len(s), 3), np.float32) if want_acc else None
    ph = np.zeros(6, np.float64)
    rc = L.ref_octree_step(P(s), len(s), G, eps2, dt, 1 if integrate else 0, P(acc), P(ph))
    if rc != 0:
        raise RuntimeError("reference octree step failed")
    return s, acc, ph


def build_tree_pipeline(stars, theta, mn=None, mx=None, offset=False):
    """Bounds, Morton, stable sort, Karras build, ranges and aggregation on the
    CPU for one instance.  Returns a dict with every intermediate."""
    x, y, z = c32(stars["x"]), c32(stars["y"]), c32(stars["z"])
    if mn is None:
        mn, mx = bounds_cube(x, y, z)
    keys = morton_n(x, y, z, mn, mx)
    skeys, perm = stable_sort_pairs(keys)
    s = np.ascontiguousarray(stars[perm])
    left, right, parent = karras_build(skeys)
    n = len(s)
    first, last = node_ranges(n, left, right) if n > 0 else (None, None)
    agg = aggregate(left, right, s["x"], s["y"], s["z"], s["mass"], theta, offset)
    return dict(mn=mn, mx=mx, keys=keys, skeys=skeys, perm=perm, sorted=s, left=left,
                right=right, parent=parent, first=first, last=last, agg=agg)
