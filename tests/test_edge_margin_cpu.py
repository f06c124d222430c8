"""The depth pass's edge pre-test keeps every pixel the exact ray/triangle test accepts: over
about 200 M pixel tests of random triangles at the renderer's scale (slivers, near and far
cameras, every orientation), the most negative raw edge distance of an accepted pixel stays
far inside the pre-test's margin (checked on the host with the kernels' own header), and the
stepped walk of small boxes keeps every accepted pixel."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "edge_margin_check.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_pretest_margin_covers_exact_test(tmp_path):
    exe = str(tmp_path / "edge_margin_check")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                        "--offload-arch=gfx950", SRC, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe, "400000"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0
    f = out.stdout.split()
    vals = dict(zip(f[0::2], f[1::2]))
    assert int(vals["accepted"]) > 10_000_000
    worst, margin = float(vals["worst"]), float(vals["margin"])
    assert worst > -margin / 20.0, (worst, margin)
    # the stepped walk of boxes up to 64 px rejects no accepted pixel, and strays from the
    # direct edge values by far less than the margin
    assert int(vals["walk_boxes"]) > 100_000
    assert int(vals["walk_missed"]) == 0
    assert float(vals["walk_dev"]) < margin / 20.0
