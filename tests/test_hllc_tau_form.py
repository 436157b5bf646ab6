"""The HLLC solver's rest-mass-free energy form equals the reference-order
solver.

The FP32 HLLC fluxes carry the frame energy as T^(a)_(0) + rho u^(a) (formed
with u^(0) - 1 = q^2 / (u^(0) + 1)), the star state's energy with
T^(0)_(dir) - rho u^(dir), the contact speed from the root 2C / (-B + sqrt(D))
and the coordinate energy flux as the spatial frame rows plus alpha times the
frame energy plus (1 - alpha) times the mass flux: no step cancels the rest
mass.  `hllc_frame_flux<DIR, double, true>` runs that form in FP64; the check
compiles a host program over it and the reference-order form
(fluxcalc2D_FT, GPU_hllc.h:3-247) and compares them on 120000 Kerr-Schild
faces with random hydrodynamic and magnetised states: within 1e-9 of
|F_rho| + |F| wherever the FP64 solver is well conditioned (a density-field
rescaling that scales the fluxes exactly moves them by less than 1e-11).  A
term lost in the form (the energy row read after the Maxwell row overwrote
the frame momentum it needs) leaves errors of order one here."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "hamr_gamma_wind_amd")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LIB = os.path.join(PKG, "libhamr_hip.so")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
@pytest.mark.skipif(not os.path.exists(LIB), reason="libhamr_hip.so not built")
def test_tau_form_hllc_equals_reference_order_hllc(tmp_path):
    exe = str(tmp_path / "check_hllc_tau")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(PKG, "csrc"),
                        "-x", "hip", "--offload-arch=gfx950", os.path.join(HERE, "hllc_tau", "check_hllc_tau.cpp"),
                        "-o", exe, "-L" + PKG, "-l:libhamr_hip.so", "-Wl,-rpath," + PKG],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
