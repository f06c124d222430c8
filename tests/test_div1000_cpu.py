"""div1000 (the FMA-based division by 1000 of the ray/triangle setup) equals the IEEE binary32
division for every non-negative float and a sample of negative ones (host check)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "div1000_check.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_div1000_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div1000_check")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                        "-mfma", "--offload-arch=gfx950", SRC, "-o", exe, "-pthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    print(out.stdout)
    assert out.returncode == 0 and "mismatches 0 negative_sample_mismatches 0" in out.stdout
