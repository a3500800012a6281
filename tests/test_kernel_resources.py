"""Register budgets of the headline kernels, from hipcc's resource-usage remarks for gfx950 (CPU
only: hipcc cross-compiles). A change that makes the Lorenz forward spill, or pushes the plain
pair adjoint past its 128-VGPR four-wave budget, costs the headline workload measurably (a
pair adjoint that spilled 24 VGPRs ran ~1.5 ms instead of ~1.15 ms at N = 2^20), so it fails
here before it reaches a GPU.
"""
import os
import re
import subprocess
import tempfile

import pytest

from gradsolve_amd import build as B

REMARK = re.compile(r"remark:\s+(?:Function Name: (\S+)|(VGPRs|AGPRs|ScratchSize \[bytes/lane\]|"
                    r"Occupancy \[waves/SIMD\]|VGPRs Spill): (\d+))")


@pytest.fixture(scope="module")
def tsit5_usage():
    gendir = tempfile.mkdtemp(prefix="gsx_ru_")
    B.write_rt_headers(gendir)
    out = subprocess.run([B.hipcc(), *B.HIPCC_FLAGS, f"-I{gendir}", "-c",
                          os.path.join(B.CSRC, "gsx_tsit5.hip"), "-o",
                          os.path.join(gendir, "t.o"), "-Rpass-analysis=kernel-resource-usage"],
                         capture_output=True, text=True, check=True)
    usage, cur = {}, None
    for m in REMARK.finditer(out.stderr):
        if m.group(1):
            cur = usage.setdefault(m.group(1), {})
        elif cur is not None:
            cur[m.group(2)] = int(m.group(3))
    return usage


def _find(usage, pattern):
    names = [k for k in usage if re.search(pattern, k)]
    assert len(names) == 1, (pattern, names)
    return usage[names[0]]


def test_headline_forward_does_not_spill(tsit5_usage):
    u = _find(tsit5_usage, r"^_ZN3gsx16tsit5_fwd_kernelINS_7FLorenzELb1EEEvNS_7FwdArgsE$")
    assert u["VGPRs Spill"] == 0 and u["ScratchSize [bytes/lane]"] == 0
    assert u["Occupancy [waves/SIMD]"] >= 4


def test_plain_pair_adjoint_holds_four_waves(tsit5_usage):
    u = _find(tsit5_usage, r"^_ZN3gsx21tsit5_adj_pair_kernelINS_7FLorenzELb0EEEvNS_7AdjArgsE$")
    assert u["VGPRs"] <= 128 and u["Occupancy [waves/SIMD]"] >= 4
    assert u["VGPRs Spill"] <= 2, u
    s = _find(tsit5_usage, r"^_ZN3gsx21tsit5_adj_pair_kernelINS_7FLorenzELb1EEEvNS_7AdjArgsE$")
    assert s["VGPRs Spill"] == 0, s
