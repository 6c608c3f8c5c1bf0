"""Python view of the engine C ABI (include/gcs.h): used by the tests, the
benchmark and the job server.  Every call goes to libgcs_amd.so."""
import ctypes

import numpy as np

from . import _lib as L


def params(**kw):
    """gcs_params with the library defaults, overridden by keyword."""
    p = L.Params()
    L.lib().gcs_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def reference_params(G, eps2, dt, global_min, global_max, **kw):
    """The semantics of cuda_physics_step (cuda_engine.cu:379-451)."""
    p = params(G=G, eps2=eps2, dt=dt, mac=L.MAC_REFERENCE, use_bounds=1,
               global_min=global_min, global_max=global_max, colours=1)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def generate(spec):
    n = spec.n_a + spec.n_b
    out = np.zeros(n, L.STAR)
    L.check(L.lib().gcs_generate(ctypes.byref(spec), L.ptr(out)))
    return out


def collision(n_per_galaxy, seed):
    """The single-collision configuration (two galaxies of n particles)."""
    s = L.CollisionSpec()
    L.lib().gcs_spec_default(n_per_galaxy, seed, ctypes.byref(s))
    return generate(s)


def random_spec(seed, k, n_min, n_max):
    s = L.CollisionSpec()
    L.lib().gcs_spec_random(seed, k, n_min, n_max, ctypes.byref(s))
    return s


class Engine:
    def __init__(self, capacity, instances=1, device=0):
        h = ctypes.c_void_p()
        L.check(L.lib().gcs_create(int(capacity), int(instances), int(device), ctypes.byref(h)))
        self._h = h
        self.capacity = int(capacity)

    def close(self):
        if self._h:
            L.lib().gcs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- state -----------------------------------------------------------
    def load(self, stars, counts=None):
        stars = np.ascontiguousarray(stars, dtype=L.STAR)
        if counts is None:
            counts = [len(stars)]
        c = np.ascontiguousarray(counts, dtype=np.int32)
        L.check(L.lib().gcs_load(self._h, L.ptr(stars), L.ptr(c), len(c)))

    def load_generated(self, specs):
        """Generates the collisions of `specs` (CollisionSpec list) on the
        device, straight into the engine (gcs_load_generated)."""
        arr = (L.CollisionSpec * len(specs))(*specs)
        L.check(L.lib().gcs_load_generated(self._h, arr, len(specs)))

    def add_generated(self, spec):
        """Adds one collision generated on the device (gcs_instance_add_generated)."""
        i = ctypes.c_int32()
        L.check(L.lib().gcs_instance_add_generated(self._h, ctypes.byref(spec), ctypes.byref(i)))
        return i.value

    def add(self, stars):
        stars = np.ascontiguousarray(stars, dtype=L.STAR)
        i = ctypes.c_int32()
        L.check(L.lib().gcs_instance_add(self._h, L.ptr(stars), len(stars), ctypes.byref(i)))
        return i.value

    def remove(self, iid):
        L.check(L.lib().gcs_instance_remove(self._h, int(iid)))

    def count(self):
        k = ctypes.c_int32()
        n = ctypes.c_int64()
        L.check(L.lib().gcs_instance_count(self._h, ctypes.byref(k), ctypes.byref(n)))
        return k.value, n.value

    def info(self, iid):
        n = ctypes.c_int32()
        off = ctypes.c_int64()
        L.check(L.lib().gcs_instance_info(self._h, int(iid), ctypes.byref(n), ctypes.byref(off)))
        return n.value, off.value

    def stars(self):
        out = np.zeros(self.count()[1], L.STAR)
        L.check(L.lib().gcs_read_stars(self._h, L.ptr(out)))
        return out

    def instance(self, iid):
        n, _ = self.info(iid)
        out = np.zeros(n, L.STAR)
        L.check(L.lib().gcs_read_instance(self._h, int(iid), L.ptr(out)))
        return out

    def checkpoint(self, iid):
        """Checkpoint of one instance (gcs_checkpoint) as bytes."""
        n, _ = self.info(iid)
        size = L.lib().gcs_checkpoint_bytes(n)
        buf = ctypes.create_string_buffer(size)
        L.check(L.lib().gcs_checkpoint(self._h, int(iid), buf, size))
        return buf.raw

    def restore(self, data):
        """Appends the instance held in a checkpoint; returns its id."""
        iid = ctypes.c_int32()
        buf = ctypes.create_string_buffer(data, len(data))
        L.check(L.lib().gcs_restore(self._h, buf, len(data), ctypes.byref(iid)))
        return iid.value

    def position_lo(self):
        out = np.zeros((self.count()[1], 3), np.float32)
        L.check(L.lib().gcs_read_position_lo(self._h, L.ptr(out)))
        return out

    def playback(self):
        out = np.zeros(self.count()[1], L.PLAYBACK_STAR)
        L.check(L.lib().gcs_read_playback(self._h, L.ptr(out)))
        return out

    def bake(self, p, frames, path):
        """gcs_bake: `frames` steps written to `path` as PlaybackStar frames;
        returns the wall seconds."""
        sec = ctypes.c_double()
        L.check(L.lib().gcs_bake(self._h, ctypes.byref(p), int(frames), str(path).encode(),
                                 ctypes.byref(sec)))
        return sec.value

    # ---- stepping --------------------------------------------------------
    def step(self, p, nsteps=1, sync=True):
        L.check(L.lib().gcs_step(self._h, ctypes.byref(p), int(nsteps)))
        if sync:
            self.sync()

    def sync(self):
        L.check(L.lib().gcs_sync(self._h))

    def compute_accels(self, p):
        L.check(L.lib().gcs_compute_accels(self._h, ctypes.byref(p)))

    def autotune(self, p, reps=3):
        """Fastest traversal settings for the current state (gcs_autotune):
        returns (params, ms per accelerations-only pass)."""
        best = L.Params()
        ms = ctypes.c_float()
        L.check(L.lib().gcs_autotune(self._h, ctypes.byref(p), int(reps), ctypes.byref(best),
                                     ctypes.byref(ms)))
        return best, ms.value

    def accels(self):
        out = np.zeros((self.count()[1], 4), np.float32)
        L.check(L.lib().gcs_read_accels(self._h, L.ptr(out)))
        return out

    def keys(self):
        n = self.count()[1]
        raw = np.zeros(n, np.uint64)
        srt = np.zeros(n, np.uint64)
        perm = np.zeros(n, np.int32)
        L.check(L.lib().gcs_read_keys(self._h, L.ptr(raw), L.ptr(srt), L.ptr(perm)))
        return raw, srt, perm

    def tree(self, iid=0):
        n, _ = self.info(iid)
        m = 2 * n - 1
        left = np.zeros(m, np.int32)
        right = np.zeros(m, np.int32)
        parent = np.zeros(m, np.int32)
        com4 = np.zeros((m, 4), np.float32)
        width = np.zeros(m, np.float32)
        bmin = np.zeros((m, 3), np.float32)
        bmax = np.zeros((m, 3), np.float32)
        L.check(L.lib().gcs_read_tree(self._h, int(iid), L.ptr(left), L.ptr(right),
                                      L.ptr(parent), L.ptr(com4), L.ptr(width), L.ptr(bmin),
                                      L.ptr(bmax)))
        return dict(left=left, right=right, parent=parent, com4=com4, width=width, bmin=bmin,
                    bmax=bmax)

    def root(self, iid=0):
        """Root record (com x, y, z, mass) of the last built tree (gcs_read_root)."""
        out = np.zeros(4, np.float32)
        L.check(L.lib().gcs_read_root(self._h, int(iid), L.ptr(out)))
        return out

    def bounds(self, iid=0):
        a = ctypes.c_float()
        b = ctypes.c_float()
        L.check(L.lib().gcs_read_bounds(self._h, int(iid), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def interactions(self):
        out = np.zeros(2, np.int64)
        L.check(L.lib().gcs_read_interactions(self._h, L.ptr(out)))
        return int(out[0]), int(out[1])

    def interaction_counts(self):
        """[per-target node, particle, monopole, local (group, node) pairs,
        target interactions of those] of the last group traversal."""
        out = np.zeros(5, np.int64)
        L.check(L.lib().gcs_read_interaction_counts(self._h, L.ptr(out)))
        return [int(v) for v in out]

    def monopole_interactions(self):
        out = np.zeros(1, np.int64)
        L.check(L.lib().gcs_read_monopole_interactions(self._h, L.ptr(out)))
        return int(out[0])

    def diagnostics(self, iid=0):
        out = np.zeros(5, np.float64)
        L.check(L.lib().gcs_read_diagnostics(self._h, int(iid), L.ptr(out)))
        return out

    def max_accel(self, iid=0):
        out = ctypes.c_float()
        L.check(L.lib().gcs_read_max_accel(self._h, int(iid), ctypes.byref(out)))
        return out.value

    def sim_time(self, iid=0):
        out = ctypes.c_double()
        L.check(L.lib().gcs_read_sim_time(self._h, int(iid), ctypes.byref(out)))
        return out.value

    def engine_info(self):
        out = np.zeros(8, np.int64)
        L.check(L.lib().gcs_engine_info(self._h, L.ptr(out)))
        keys = ["num_cu", "trav_blocks", "trav_waves", "trav_slot_bytes", "capacity",
                "instance_capacity", "groups", "tiles"]
        return dict(zip(keys, (int(x) for x in out)))

    def stage_ms(self):
        out = np.zeros(7, np.float32)
        L.check(L.lib().gcs_read_stage_ms(self._h, L.ptr(out)))
        return out


def autotune_cache_clear():
    L.check(L.lib().gcs_autotune_cache_clear())


def autotune_cache_save(path):
    L.check(L.lib().gcs_autotune_cache_save(path.encode()))


def autotune_cache_load(path):
    L.check(L.lib().gcs_autotune_cache_load(path.encode()))


def morton_keys(x, y, z, mn, mx):
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    out = np.zeros(len(x), np.uint64)
    L.check(L.lib().gcs_morton_keys(L.ptr(x), L.ptr(y), L.ptr(z), len(x), mn, mx, L.ptr(out)))
    return out


def sort_pairs(keys):
    keys = np.ascontiguousarray(keys, np.uint64)
    ko = np.zeros(len(keys), np.uint64)
    po = np.zeros(len(keys), np.int32)
    L.check(L.lib().gcs_sort_pairs(L.ptr(keys), len(keys), L.ptr(ko), L.ptr(po)))
    return ko, po


def karras_tree(sorted_keys):
    k = np.ascontiguousarray(sorted_keys, np.uint64)
    n = len(k)
    left = np.zeros(2 * n - 1, np.int32)
    right = np.zeros(2 * n - 1, np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    L.check(L.lib().gcs_karras_tree(L.ptr(k), n, L.ptr(left), L.ptr(right), L.ptr(parent)))
    return left, right, parent
