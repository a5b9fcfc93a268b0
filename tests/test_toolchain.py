"""Guard against a gfx950 code-generation fault the library's build flags avoid.

A three-way selection on a workgroup-uniform value (blockIdx.y, or a kernel
argument) with the last case as the fall-through, for example

    if (e == 0) p = f + sz; else if (e == 1) p = f + sy; else p = f + 1;

is structurized by hipcc (ROCm 7.2) at -O3 so that the final case's address
is never computed: the load then goes through registers that hold another
value and the kernel takes a memory access fault.  The library is built with
``-mllvm -structurizecfg-skip-uniform-regions=true``, under which uniform
branches stay plain scalar branches.  This test compiles such a kernel to
gfx950 assembly with the library's flags and checks that no vector register
reaches the load undefined.
"""
import os
import re
import shutil
import subprocess

import pytest

from lowmachmixtures_amd import build as B

SRC = r"""
#include <hip/hip_runtime.h>
struct P3 { double *c[3]; };
struct G { long s[3]; };
__global__ void probe_uniform_select(G g, P3 fi, P3 co, long n) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int e = (int)blockIdx.y;
  const double *f = fi.c[e];
  double v;
  if (e == 0) v = f[t] + f[t + g.s[2]];
  else if (e == 1) v = f[t] + f[t + g.s[1]];
  else v = f[t] + f[t + 1];
  co.c[e][t] = v;
}
"""


def compile_asm(tmp_path, flags):
    src = tmp_path / "probe.hip"
    src.write_text(SRC)
    out = tmp_path / "probe.s"
    cmd = [B.hipcc()] + [f for f in flags if f not in ("-shared", "-fPIC")] + [
        "--cuda-device-only", "-S", str(src), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True)
    return out.read_text()


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="hipcc not available")
def test_uniform_three_way_select_is_fully_computed(tmp_path):
    asm = compile_asm(tmp_path, B.flags())
    m = re.search(r"^_Z\w*probe_uniform_select\w*:(.*?)\.Lfunc_end", asm, re.S | re.M)
    assert m, "kernel not found in the assembly"
    body = m.group(1)
    assert "implicit-def: $vgpr" not in body
    # all three address forms are present: + s[2], + s[1] and + 1 element (8 bytes)
    assert len(re.findall(r"global_load_dwordx2", body)) >= 2
