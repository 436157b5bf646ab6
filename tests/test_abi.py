"""The C ABI: libhamr_hip.so loads and exports every function include/hamr_hip.h
declares (no device calls), the header is plain C, and host-only entry points
reject bad descriptors without touching a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

import hamr_gamma_wind_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hamr_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hamr_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("hamr_create", "hamr_destroy", "hamr_fluxcalc", "hamr_block_dt", "hamr_flux_correct", "hamr_ct_emf", "hamr_ct_bound",
                 "hamr_ct_update", "hamr_update", "hamr_boundprim", "hamr_flags", "hamr_step", "hamr_upload",
                 "hamr_download", "hamr_upload_metric", "hamr_rdump_block_write", "hamr_rdump_block_read",
                 "hamr_rdump_block_size", "hamr_rdump_offsets", "hamr_rdump_write_dev", "hamr_rdump_read_dev"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = H.load_library()
    host = H.load_host_library()
    # hamr_host_* live in libhamr_host.so (the C host over the ABI), the rest in libhamr_hip.so
    missing = [n for n in declared_functions() if not hasattr(host if n.startswith("hamr_host_") else lib, n)]
    assert not missing, missing


def test_header_is_plain_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "hamr_hip.h"\nint main(void){ hamr_desc d; (void)d; return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.dirname(HEADER), "-c", str(src),
                        "-o", str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "s.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "hamr_hip.h"\n'
                   'int main(void){ printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(hamr_desc), offsetof(hamr_desc, minkowski),'
                   ' sizeof(hamr_torus_desc), offsetof(hamr_torus_desc, seed), sizeof(hamr_amr_desc),'
                   ' offsetof(hamr_amr_desc, p_mid)); return 0; }\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(H.Desc)
    assert int(out[1]) == H.Desc.minkowski.offset
    assert int(out[2]) == C.sizeof(H.TorusDesc)
    assert int(out[3]) == H.TorusDesc.seed.offset
    assert int(out[4]) == C.sizeof(H.AmrDescC)
    assert int(out[5]) == H.AmrDescC.p_mid.offset


def test_host_entry_points_reject_bad_descriptors():
    bad = H.make_desc((4, 16, 16), (1, 1, 1))
    with pytest.raises(H.HamrError):
        H.desc_layout(bad)
    lay = H.desc_layout(H.make_desc((32, 32, 32), (8, 4, 4)))
    assert lay["D"] == (38, 38, 38) and lay["ncell"] == 38 ** 3 and lay["ksize"] % 64 == 0
    assert lay["nblk"] == 128 and lay["ngeo"] == 32
