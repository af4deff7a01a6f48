"""Register budgets of the hot kernels (CPU: compiles for gfx950 with hipcc, runs nothing on a GPU).

The stencil kernels are latency-bound: a kernel that grows past 128 VGPRs holds one 8-wave workgroup per CU
instead of two, and ran 65 % slower at the 64-simulation batch when that happened unnoticed. These checks
compile the sources with the library's flags and read the compiler's resource report."""
import os
import re
import shutil
import subprocess

import pytest

from beamz_amd import build as B

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")

_cache = {}


def resources(src):
    """{mangled kernel name: (VGPRs, AGPRs, scratch bytes per lane)} from hipcc's resource report."""
    if src in _cache:
        return _cache[src]
    out = os.path.join(B.BUILD, "resources_" + src.replace(".hip", ".o"))
    os.makedirs(B.BUILD, exist_ok=True)
    r = subprocess.run([HIPCC, *B.HIP_FLAGS, "-c", os.path.join(B.CSRC, src), "-o", out,
                        "-Rpass-analysis=kernel-resource-usage"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    res, name = {}, None
    for line in r.stderr.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            res[name] = [None, 0, 0]
            continue
        for key, idx in (("VGPRs: ", 0), ("AGPRs: ", 1), ("ScratchSize [bytes/lane]: ", 2)):
            m = re.search(re.escape(key) + r"(\d+)", line)
            if m and name:
                res[name][idx] = int(m.group(1))
    _cache[src] = {k: tuple(v) for k, v in res.items()}
    return _cache[src]


def pick(res, pattern):
    hits = {k: v for k, v in res.items() if re.search(pattern, k)}
    assert hits, pattern
    return hits


@pytest.mark.parametrize("src,pattern,budget", [
    # K-step 2D kernel, default shapes (8 x 8 and 4 x 8), without the multi-source lookup
    ("yee2d_fused.hip", r"k_yee2d_tbkILi8ELi8ELb0E", 128),
    ("yee2d_fused.hip", r"k_yee2d_tbkILi4ELi8ELb0E", 128),
    # two-step 3D kernel, default shape (2 rows x 8 waves)
    ("yee3d_tb2.hip", r"k_yee3d_tb2ILi2ELi8E", 128),
    # K-step reverse sweep of the adjoint, default shapes
    ("adjoint2d.hip", r"k_adj2d_tbkILi8ELi8E", 128),
    ("adjoint2d.hip", r"k_adj2d_tbkILi4ELi8E", 128),
])
def test_hot_kernels_keep_two_workgroups_per_cu(src, pattern, budget):
    for name, (vgpr, agpr, scratch) in pick(resources(src), pattern).items():
        assert vgpr + agpr <= budget, (name, vgpr, agpr)
        assert scratch == 0, (name, scratch)
