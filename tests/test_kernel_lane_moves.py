"""The xyz sweep kernels' DPP lane moves run with every lane of the wave
active. A move that sits in a divergent branch (for instance a move written
as an operand of ?:, which C++ evaluates on one side only) reads 0 on the
lanes whose source lane is switched off; a FP64 edge term then divided 0 by
0. This test compiles sweep_kernels.hip for gfx950 to assembly and checks
that no v_*_dpp instruction of the sweep kernels is issued between an
s_and_saveexec (which narrows exec for a branch) and the s_or_b64 exec that
restores it. No GPU is needed; it is skipped when hipcc is absent."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def dpp_under_narrowed_exec(asm, kernel_filter):
    """{kernel symbol: count of DPP instructions issued while a saveexec
    branch region is open}."""
    name, depth, bad = None, 0, {}
    for line in asm.split("\n"):
        m = re.match(r"^(_Z\S+):", line)
        if m:
            name, depth = m.group(1), 0
            continue
        t = line.strip()
        if name is None or not t or t.startswith((".", ";")):
            continue
        if t.startswith("s_and_saveexec"):
            depth += 1
        elif t.startswith("s_or_b64 exec, exec"):
            depth = max(0, depth - 1)
        elif t.startswith("v_") and "dpp" in t and depth > 0 and kernel_filter(name):
            bad[name] = bad.get(name, 0) + 1
    return bad


def test_scanner_flags_a_move_inside_a_branch(tmp_path):
    """The check itself: a kernel whose move is an operand of ?: is flagged,
    the same kernel with the move hoisted is not."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    src = r'''
#include <hip/hip_runtime.h>
__device__ __forceinline__ double mv(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x39, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x39, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__global__ void branchy(const double* a, double* o) {
  const double v = a[threadIdx.x];
  o[threadIdx.x] = v > 0 ? v : mv(v);
}
__global__ void hoisted(const double* a, double* o) {
  const double v = a[threadIdx.x];
  const double n = mv(v);
  o[threadIdx.x] = v > 0 ? v : n;
}
'''
    f = tmp_path / "probe.hip"
    f.write_text(src)
    out = tmp_path / "probe.s"
    subprocess.run([HIPCC, "-O3", "--offload-arch=gfx950", "--cuda-device-only", "-S", str(f),
                    "-o", str(out)], check=True, capture_output=True)
    bad = dpp_under_narrowed_exec(out.read_text(), lambda n: True)
    assert any("branchy" in k for k in bad), bad
    assert not any("hoisted" in k for k in bad), bad


def test_sweep_kernels_move_with_every_lane_active(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    csrc = os.path.join(ROOT, "umt_amd", "csrc")
    out = tmp_path / "sweep_kernels.s"
    subprocess.run([HIPCC, "-O3", "--offload-arch=gfx950", "-std=c++17", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S",
                    os.path.join(csrc, "sweep_kernels.hip"), "-o", str(out)],
                   check=True, capture_output=True)
    asm = out.read_text()
    kernels = [n for n in re.findall(r"^(_Z\S+):", asm, re.M) if "sweep" in n]
    assert any("ucb_sweep_level" in n for n in kernels)
    bad = dpp_under_narrowed_exec(asm, lambda n: "sweep" in n)
    assert not bad, bad
