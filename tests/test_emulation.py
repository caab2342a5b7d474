"""CPU emulation of device code, the checks behind its bit-for-bit claims and
its memory safety:
  - the one-pass smoother (mg_tb.hpp) against full-grid red-black half-sweeps;
  - the one-pass residual + restriction (mg_rr.hpp, 3D streaming and 2D)
    against the residual followed by the restriction;
  - the advection gathers (fluid_device.hpp) against the C oracle's advect /
    advect_mac, 2D and 3D, f32 and f64, one- and two-element rows,
    displacements clamped at every wall;
  - the red-black operator over runs of rows (2D) and planes (3D) against one
    row or plane per call (mg_device.hpp).
The device code is compiled as host code (hipcc, host-side AddressSanitizer,
exact-size arrays) and run one workgroup or one element at a time, each phase
over all threads before the next, on ragged shapes, every kind of boundary and
several tile and chunk sizes. No GPU is used; the same arithmetic runs on the
GPU in tests/test_gpu_smoother.py and tests/test_gpu_parity.py.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "panopaea_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ORACLE_C = os.path.join(ROOT, "oracle", "oracle.c")

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")


@pytest.mark.parametrize("name", ["smoother_emu", "rr_emu", "adv_emu", "op_emu"])
def test_emulation(name, tmp_path):
    src = os.path.join(HERE, "emu", name + ".cpp")
    exe = str(tmp_path / name)
    extra = []
    if name == "adv_emu":  # checked against the C oracle, linked in as an object
        obj = str(tmp_path / "oracle.o")
        c = subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-c", ORACLE_C,
                            "-o", obj], capture_output=True, text=True, timeout=300)
        assert c.returncode == 0, c.stderr[-4000:]
        extra = ["-x", "none", obj]
    cmd = [HIPCC, "-O1", "-std=c++17", "-ffp-contract=off", "--offload-arch=gfx950",
           "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
           "-I", CSRC, "-I", INCLUDE, src, "-o", exe] + extra
    b = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "ERROR: AddressSanitizer" not in r.stderr
