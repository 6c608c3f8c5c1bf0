"""The bucket sort's indexing on the host: tests/bucket_sort_host.cpp runs the
kernels' helper functions (gcs_sort.hpp) in host loops that mirror each
kernel and checks every instance against std::stable_sort -- nearly sorted,
widely displaced, random and equal keys (buckets above the LDS size, merged
in chunks) and a forest at the bucket edges.  No GPU: built with hipcc for
the host."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "galaxy-collision-simulator_amd", "csrc")


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(hipcc() is None, reason="hipcc not available")
def test_bucket_sort_host_emulation(tmp_path):
    exe = str(tmp_path / "bucket_sort_host")
    subprocess.run([hipcc(), "-O2", "-std=c++17", "-I", CSRC, "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "bucket_sort_host.cpp"), "-o", exe],
                   check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=300)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout
    assert "chunked=0" not in r.stdout.split("random order")[1].splitlines()[0]
