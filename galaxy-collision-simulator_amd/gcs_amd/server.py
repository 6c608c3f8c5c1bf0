"""Python view of the served form (gcs_server_* in include/gcs.h)."""
import ctypes

import numpy as np

from . import _lib as L
from .engine import params as make_params


def poisson_jobs(seed, rate, duration, n_min, n_max, steps_min, steps_max):
    """Job schedule: Poisson arrivals at `rate` per second over `duration`
    seconds, N log-uniform in [n_min, n_max] (even), steps uniform."""
    r = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    k = 0
    while True:
        t += r.exponential(1.0 / rate)
        if t >= duration:
            break
        n = int(round(np.exp(r.uniform(np.log(n_min), np.log(n_max)))))
        n = max(n_min, min(n_max, n)) // 2 * 2
        jobs.append((k, n, int(r.integers(steps_min, steps_max + 1)), t))
        k += 1
    return jobs


class Server:
    def __init__(self, hbm_budget_bytes, max_jobs=4096, steps_per_round=25, gen_threads=8,
                 keep_final=False, max_resident=0, **param_kw):
        c = L.ServerConfig()
        c.hbm_budget_bytes = int(hbm_budget_bytes)
        c.max_resident = int(max_resident)
        c.max_jobs = int(max_jobs)
        c.steps_per_round = int(steps_per_round)
        c.gen_threads = int(gen_threads)
        c.keep_final = 1 if keep_final else 0
        c.params = make_params(**param_kw)
        h = ctypes.c_void_p()
        L.check(L.lib().gcs_server_create(ctypes.byref(c), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            L.lib().gcs_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, jobs, seed):
        arr = (L.Job * len(jobs))()
        for i, (jid, n, steps, t) in enumerate(jobs):
            arr[i].id, arr[i].n, arr[i].steps, arr[i].seed, arr[i].arrival_s = jid, n, steps, seed, t
        L.check(L.lib().gcs_server_submit(self._h, arr, len(jobs)))

    def run(self, duration=0.0):
        L.check(L.lib().gcs_server_run(self._h, float(duration)))

    def results(self):
        n = ctypes.c_int32()
        L.check(L.lib().gcs_server_results(self._h, None, 0, ctypes.byref(n)))
        arr = (L.JobResult * max(n.value, 1))()
        L.check(L.lib().gcs_server_results(self._h, arr, n.value, ctypes.byref(n)))
        return [dict(id=a.id, n=a.n, steps=a.steps, arrival=a.arrival_s, admit=a.admit_s,
                     finish=a.finish_s) for a in arr[:n.value]]

    def stats(self):
        out = np.zeros(11, np.float64)
        L.check(L.lib().gcs_server_stats(self._h, L.ptr(out), len(out)))
        keys = ["particle_steps", "run_s", "finished", "admitted", "waiting", "mean_resident",
                "rounds", "resident_cap", "stepping_s", "full_particle_steps", "full_s"]
        return dict(zip(keys, (float(x) for x in out)))

    def final(self, jid, n):
        out = np.zeros(n, L.STAR)
        L.check(L.lib().gcs_server_final(self._h, int(jid), L.ptr(out), int(n)))
        return out
