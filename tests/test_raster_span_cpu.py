"""The large-face span walk enumerates a superset of the pixels the edge pre-test passes
(checked on the host with the same header the kernels use)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "raster_span_check.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_row_spans_cover_every_pretest_pixel(tmp_path):
    exe = str(tmp_path / "raster_span_check")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
                        "--offload-arch=gfx950", SRC, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout[-2000:]
    fields = dict(zip(out.stdout.split()[-10::2], map(int, out.stdout.split()[-9::2])))
    assert fields["missed"] == 0 and fields["passed"] > 0
    assert fields["span_px"] < 0.6 * fields["box_px"]      # the spans are much tighter than the boxes
