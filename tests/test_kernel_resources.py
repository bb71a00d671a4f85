"""Compiled resources of the hot kernels (CPU: hipcc cross-compiles gfx950 here).

The chained level kernel and the multi-job local-index kernels carry launch bounds that were chosen
by measurement; these checks read hipcc's resource report for the library's own compile flags and
fail when a change to those kernels loses the occupancy the bounds were set for, or spills more.
"""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

import util

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# kernel (a fragment of its mangled name) -> (minimum waves per SIMD, maximum scratch bytes per lane)
EXPECT = {
    "enum.hip": {
        "k_level_flat_chain_multiILi0E": (6, 0),
        "k_level_flat_chain_multiILi1E": (6, 16),
    },
    "index.hip": {
        "k_local_queue_multi": (8, 0),
        "k_local_check_multi": (7, 0),
    },
}


def library_hipflags():
    """The HIPFLAGS line of the library's Makefile, for gfx950."""
    for line in open(os.path.join(util.PKG, "Makefile")):
        if line.startswith("HIPFLAGS"):
            return line.split(":=", 1)[1].replace("$(ARCH)", "gfx950").split()
    raise AssertionError("HIPFLAGS missing from the Makefile")


def resource_report(text):
    """{mangled kernel name: {"VGPRs": n, "VGPRs Spill": n, "ScratchSize": n, "Occupancy": n}}"""
    out, cur = {}, None
    for line in text.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            cur = out.setdefault(m.group(1), {})
            continue
        m = re.search(r"remark:\s+(VGPRs Spill|VGPRs|ScratchSize|Occupancy)\b[^:]*:\s*(\d+)", line)
        if m and cur is not None:
            cur[m.group(1)] = int(m.group(2))
    return out


@pytest.mark.skipif(not (os.path.exists(HIPCC) or shutil.which("hipcc")), reason="hipcc not installed")
def test_hot_kernels_keep_their_occupancy():
    flags = [f for f in library_hipflags() if f != "-Wall"]
    with tempfile.TemporaryDirectory() as tmp:
        procs = {}
        for src in EXPECT:
            cmd = [HIPCC] + flags + ["--cuda-device-only", "-c", os.path.join(util.PKG, "csrc", "device", src),
                                     "-o", os.path.join(tmp, src + ".o"), "-Rpass-analysis=kernel-resource-usage"]
            procs[src] = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        reports = {}
        for src, p in procs.items():
            text, _ = p.communicate(timeout=600)
            assert p.returncode == 0, text[-2000:]
            reports[src] = resource_report(text)
    for src, kernels in EXPECT.items():
        for frag, (waves, scratch) in kernels.items():
            hits = [r for name, r in reports[src].items() if frag in name]
            assert len(hits) == 1, (src, frag, sorted(reports[src]))
            r = hits[0]
            assert r["Occupancy"] >= waves, (frag, r)
            assert r["ScratchSize"] <= scratch, (frag, r)
            if scratch == 0:
                assert r.get("VGPRs Spill", 0) == 0, (frag, r)
