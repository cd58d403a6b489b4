"""The MIOpen baseline program (convstencil_amd/bin/miopen_stencil, the counterpart of the
reference's cuDNN programs, src/cudnn/*.cu) on the GPU: it runs, reports the reference's
Time / GStencil/s lines, and its result matches ConvStencil's FP64 result to the bar of the
precision MIOpen ran in (FP64: 1e-12 relative L2; FP32: 1e-5).
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "convstencil_amd", "bin",
                   "miopen_stencil")


@pytest.mark.parametrize("args", [["box2d1r", "200", "300", "4"], ["1d1r", "5000", "6"],
                                  ["box3d1r", "20", "24", "28", "3"], ["star2d3r", "100", "90", "2"]])
def test_miopen_baseline_matches(gpu, args):
    print("case", args, flush=True)
    p = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=600)
    print(p.stdout, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert re.search(r"Time = \d+\[ms\]\nGStencil/s = [0-9.]+\n", p.stdout), p.stdout
    m = re.search(r"MIOpen\(\dD\) \S+ (fp64|fp32), algorithm (\S+)", p.stdout)
    l2 = float(re.search(r"rel_l2 vs ConvStencil fp64 = ([0-9.e+-]+)", p.stdout).group(1))
    assert l2 <= (1e-12 if m.group(1) == "fp64" else 1e-5), p.stdout
