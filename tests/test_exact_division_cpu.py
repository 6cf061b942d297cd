"""The shadow filter divides the luminance error by 3 through the reciprocal
(vpt_svgf.hip, div3_unit): this checks that sequence against the IEEE division
for every float in [0, 2] with a C program (fmaf from the C library, which is
exact, as the GPU's v_fma_f32 is)."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div3_unit_equals_division_on_0_2(tmp_path):
    exe = tmp_path / "div3_check"
    # -ffp-contract=off: x * y and the divisions round as written
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", str(HERE / "native" / "div3_check.c"), "-o", str(exe), "-lm"],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("0 mismatches")
