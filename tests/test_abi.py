"""The C-ABI library loads and exports every function the public headers declare;
host-side validation paths return error codes instead of exiting (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "fletcher", h)
           for h in ("driver.h", "fl_hip.h", "fl_model.h", "fl_run.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        with open(h) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            before = text[:m.start()].rstrip()
            if name in ("if", "sizeof", "defined") or before.endswith("#define"):
                continue
            # a declaration: preceded by a type token at statement start
            stmt = before[max(before.rfind(";"), before.rfind("}"), before.rfind("{")) + 1:]
            if re.search(r"(void|int|float|double|char|long|int32_t|fl_rsf_t|const)\s*\**\s*$", stmt.strip() + " ") \
                    or re.search(r"\b(void|int|float|double|long long|int32_t)\s*\**$", stmt.strip()):
                names.add(name)
    return sorted(n for n in names if n.startswith(("DRIVER_", "fl_")))


def test_headers_declare_the_reference_driver_api():
    names = declared_functions()
    for n in ("DRIVER_Initialize", "DRIVER_Finalize", "DRIVER_Propagate", "DRIVER_Update_pointers",
              "DRIVER_InsertSource"):
        assert n in names
    assert len(names) > 50


def test_library_exports_every_declared_function(fl):
    L = fl.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_cli_is_linked_against_the_library(fl):
    import subprocess
    out = subprocess.run(["ldd", fl.CLI_PATH], capture_output=True, text=True).stdout
    assert "libfletcher_amd.so" in out and "not found" not in out


def test_error_paths_return_codes(fl):
    L = fl.lib()
    L.fl_clear_error()
    assert L.fl_set_variant(7) == -1 and b"variant" in L.fl_last_error()
    assert L.fl_set_precision(9) == -1
    assert L.fl_set_tiles(5, 0) == -1
    assert L.fl_set_tiles(8, -3) == -1
    L.fl_clear_error()
    # out-of-order calls: recorded, never exit()
    L.DRIVER_Propagate(24, 24, 24, 4, 12.5, 12.5, 12.5, 0.001, 1, None, None, None, None)
    assert L.fl_status() == -2 and b"before DRIVER_Initialize" in L.fl_last_error()
    L.fl_clear_error()
    L.DRIVER_InsertSource(0.001, 0, 10, None, None, 1.0)
    assert L.fl_status() == -2
    L.DRIVER_Update_pointers(24, 24, 24, None)
    assert L.fl_status() == -2
    assert L.fl_get_field(1, np.zeros(1, np.float32)) == -2
    assert L.fl_run_steps(0, 3, 10, np.zeros(3, np.float32), 0) == -2
    assert L.fl_batch_run(0, 3, np.zeros(3, np.float32), 0, 0, 0) == -2
    L.fl_clear_error()
    assert L.fl_status() == 0


def test_initialize_rejects_bad_shapes_before_touching_the_device(fl):
    L = fl.lib()
    z = np.zeros(8, np.float32)
    L.fl_clear_error()
    L.DRIVER_Initialize(24, 24, 24, 3, 12.5, 12.5, 12.5, 0.001, z, z, z, z, z, z, None, None, None, None)
    assert L.fl_status() == -1 and b"bord" in L.fl_last_error()
    L.DRIVER_Initialize(8, 24, 24, 4, 12.5, 12.5, 12.5, 0.001, z, z, z, z, z, z, None, None, None, None)
    assert L.fl_status() == -1
    L.DRIVER_Initialize(24, 24, 24, 4, 12.5, 12.5, 12.5, -0.001, z, z, z, z, z, z, None, None, None, None)
    assert L.fl_status() == -1
    assert L.fl_batch_create(0, None, 4, 12.5, 12.5, 12.5, 0.001, 0) == -1
    L.fl_clear_error()
