"""The C ABI library loads and exports every function include/diffbmp_amd.h
declares (no GPU work: only symbol lookup and argument validation paths that
return before touching the device)."""
import ctypes
import os
import re

import pytest

from diffbmp_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "diffbmp_amd.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dbmp_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("dbmp_init", "dbmp_destroy", "dbmp_bin", "dbmp_forward", "dbmp_backward",
                 "dbmp_template_extents", "dbmp_loss_mse", "dbmp_adam_step", "dbmp_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(declared()) <= set(_native.EXPORTS), set(declared()) - set(_native.EXPORTS)


def test_version_and_argument_errors_without_device():
    L = _native.lib()
    assert L.dbmp_version() >= 100
    h = ctypes.c_void_p()
    # invalid arguments are rejected before any HIP call
    assert L.dbmp_init(ctypes.byref(h), 64, 64, 24, 0.0, 1.0, 151, 0) == 1
    assert b"multiple of 16" in L.dbmp_last_error()
    assert L.dbmp_init(ctypes.byref(h), 0, 64, 16, 0.0, 1.0, 151, 0) == 1
    assert L.dbmp_init(None, 64, 64, 16, 0.0, 1.0, 151, 0) == 1
    assert L.dbmp_init(ctypes.byref(h), 64, 64, 16, 0.0, -0.5, 151, 0) == 1
    assert b"alpha_upper_bound" in L.dbmp_last_error()
    assert L.dbmp_init(ctypes.byref(h), 64, 64, 16, 0.0, float("nan"), 151, 0) == 1
    assert L.dbmp_forward(None, None, None, None, 0.0, None, None, None, None, None) == 1
    assert L.dbmp_destroy(None) == 0
    b = ctypes.c_void_p()
    assert L.dbmp_batch_create(ctypes.byref(b), 7, 0.9, 0.999, 1e-8) == 1
    assert b"storage dtype" in L.dbmp_last_error()
    assert L.dbmp_batch_step(None, 0, None) == 1
    assert L.dbmp_batch_destroy(None) == 0


def test_struct_layouts_match_header():
    # field counts of the ctypes mirrors against the header's struct bodies
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for cname, py in (("dbmp_prims", _native.Prims), ("dbmp_bitmaps", _native.Bitmaps),
                      ("dbmp_tile_lists", _native.TileLists), ("dbmp_adam_group", _native.AdamGroup),
                      ("dbmp_adam", _native.Adam), ("dbmp_job", _native.Job),
                      ("dbmp_layer_prims", _native.LayerPrims)):
        body = re.search(r"typedef struct \{([^{}]*)\}\s*" + cname + ";", text).group(1)
        decls = [d for d in body.split(";") if d.strip()]
        n = sum(len(d.split(",")) for d in decls)
        # arrays (lo[2], groups[5]) are one ctypes field each
        assert n == len(py._fields_), (cname, n, len(py._fields_))


def test_struct_offsets_match_c_compiler(tmp_path):
    """Every ctypes mirror has the C compiler's size and field offsets."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"dbmp_prims": _native.Prims, "dbmp_bitmaps": _native.Bitmaps,
               "dbmp_tile_lists": _native.TileLists, "dbmp_adam_group": _native.AdamGroup,
               "dbmp_adam": _native.Adam, "dbmp_job": _native.Job,
               "dbmp_layer_prims": _native.LayerPrims}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = structs[cname]
        want = ctypes.sizeof(py) if field == "size" else getattr(py, field).offset
        assert int(val) == want, (cname, field, int(val), want)
