"""The host-side set-up of the library and the oracle under AddressSanitizer
and UndefinedBehaviorSanitizer: tests/sanitize/driver.c builds the block
tables and the reference's block numbering for five block counts, the metric,
tetrads and torus state of a one-block grid, and the 90-block three-level set
with its exchange tables, geometry and state (hamr_setup.cpp, hamr_amr.cpp and
the host entry points of hamr_capi.hip, compiled for the host only), and runs
the oracle on them (two HLL steps and one HLLC step, the inversion cascade with
distorted energies, restart and dump records, one multi-level step).  Any
out-of-bounds access, use after free, leak or undefined operation aborts the
run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs make and hipcc")
def test_host_setup_and_oracle_clean_under_asan_ubsan():
    r = subprocess.run(["make", "-s", "-j8", "-C", os.path.join(HERE, "sanitize"), "run"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=900)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32" not in l)
    print(out[-3000:])
    assert r.returncode == 0, out[-6000:]
    assert "sanitized run complete" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
