"""MI355X (gfx950) per-frame N-body step of the galaxy-collision simulator.

The compute path is libgcs_amd.so (HIP kernels + C ABI, include/gcs.h); this
package is its Python view, used by the tests, the benchmark and the job
server."""
from ._lib import (STAR, PLAYBACK_STAR, MAC_GROUP, MAC_REFERENCE, PREC_COMPENSATED, PREC_PLAIN,
                   GcsError, LIB_PATH, lib, NOALLOC_MAX_KEYS, ERR_INVALID, ERR_CAPACITY)
from .engine import (Engine, autotune_cache_clear, autotune_cache_load, autotune_cache_save,
                     collision, generate, karras_tree, morton_keys, params, random_spec,
                     reference_params, sort_pairs)
from .server import Server, poisson_jobs

__all__ = ["STAR", "PLAYBACK_STAR", "MAC_GROUP", "MAC_REFERENCE", "PREC_COMPENSATED",
           "PREC_PLAIN", "GcsError", "LIB_PATH", "lib", "NOALLOC_MAX_KEYS", "ERR_INVALID",
           "ERR_CAPACITY", "Engine", "autotune_cache_clear", "autotune_cache_load",
           "autotune_cache_save", "collision", "generate",
           "karras_tree", "morton_keys", "params", "random_spec", "reference_params",
           "sort_pairs", "Server", "poisson_jobs"]
