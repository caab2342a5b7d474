"""CPU emulation of the fused multigrid kernels, the checks behind their
bit-for-bit claims: the one-pass smoother (mg_tb.hpp) against full-grid
red-black half-sweeps, and the one-pass residual + restriction (mg_rr.hpp,
3D streaming and 2D) against the residual followed by the restriction. The
device structs are compiled as host code (hipcc, host-side AddressSanitizer)
and run one workgroup at a time, each phase over all threads before the next,
on ragged shapes, every kind of boundary and several tile and chunk sizes. No
GPU is used; the same arithmetic runs on the GPU in tests/test_gpu_smoother.py.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "panopaea_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")


@pytest.mark.parametrize("name", ["smoother_emu", "rr_emu"])
def test_emulation(name, tmp_path):
    src = os.path.join(HERE, "emu", name + ".cpp")
    exe = str(tmp_path / name)
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
           "-I", CSRC, "-I", INCLUDE, src, "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stderr
