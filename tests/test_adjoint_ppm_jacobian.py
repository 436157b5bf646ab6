"""The adjoint's closed-form PPM Jacobian equals its forward-mode form.

`para_jacobian` (hamr_adjoint.h) forms d(l, r)/d(x1..x5) of one cell's
limited parabola along the active branches of the MC slopes and the CW84
limiter in closed form; `para_jacobian_dual` carries five tangents through
para (GPU_interp.h:60-111) in forward mode.  The check compiles a host
program over both and compares them on 2e6 stencils (random, near-linear,
spiked and with ties), values and tangents alike."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_closed_form_ppm_jacobian_equals_forward_mode(tmp_path):
    exe = str(tmp_path / "check_ppm_jacobian")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "hamr_gamma_wind_amd", "csrc"),
                        "-x", "hip", "--offload-arch=gfx950", os.path.join(HERE, "adjoint_ppm", "check_ppm_jacobian.cpp"),
                        "-o", exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
