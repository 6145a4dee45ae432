"""The C ABI library loads, exports every entry point the header declares,
and its structs have the layout the ctypes binding assumes (no GPU calls)."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from elastoplasm import hip

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "elastoplasm.h"


def declared():
    txt = HEADER.read_text()
    return sorted(set(re.findall(r"\b(ep_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = hip.load()
    names = declared()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(hip.EXPORTS) | {"ep_version"}
    assert lib.ep_version() == int(re.search(r"EP_ABI_VERSION (\d+)", HEADER.read_text()).group(1))


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "elastoplasm.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(ep_config), "
        "offsetof(ep_config, Del), offsetof(ep_config, surrogate_tol), sizeof(ep_particles_view), "
        "sizeof(ep_step_params), sizeof(ep_step_result), sizeof(ep_stats), sizeof(ep_grid_view), "
        "sizeof(ep_grid_fields));return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), "-o", str(exe), str(src)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(hip.EpConfig), hip.EpConfig.Del.offset, hip.EpConfig.surrogate_tol.offset,
            C.sizeof(hip.EpParticlesView), C.sizeof(hip.EpStepParams), C.sizeof(hip.EpStepResult),
            C.sizeof(hip.EpStats), C.sizeof(hip.EpGridView), C.sizeof(hip.EpGridFields)]
    assert got == want


def test_config_errors_without_gpu():
    """ep_create validates the configuration before any HIP call and maps
    the reference's setup errors onto EP_E_CONFIG / EP_E_ARG."""
    lib = hip.load()
    cfg = hip.EpConfig()
    cfg.dim = 4
    h = C.c_void_p()
    assert lib.ep_create(C.byref(cfg), C.byref(h)) == -2
    assert b"dim" in lib.ep_last_error()
    cfg.dim, cfg.basis = 2, 7
    assert lib.ep_create(C.byref(cfg), C.byref(h)) == -2
    assert b"basis" in lib.ep_last_error()
    cfg.basis, cfg.plast = 0, 9
    assert lib.ep_create(C.byref(cfg), C.byref(h)) == -2
    assert b"InvalidReturnMapping" in lib.ep_last_error()
    cfg.plast, cfg.nmp = 1, 0
    assert lib.ep_create(C.byref(cfg), C.byref(h)) == -1
    assert lib.ep_create(None, C.byref(h)) == -1
    # finite deformation with a Del that is not the isotropic form
    cfg.nmp = 10
    for i in range(2):
        cfg.nel[i], cfg.nno[i], cfg.h[i] = 4, 5, 1.0
    E, nu = 1.0e6, 0.3
    f = E / ((1 + nu) * (1 - 2 * nu))
    Del = [f * (1 - nu), f * nu, 0.0, f * nu, f * (1 - nu), 0.0, 0.0, 0.0, 1.5 * f * (1 - 2 * nu) / 2]
    for k, v in enumerate(Del):
        cfg.Del[k] = v
    assert lib.ep_create(C.byref(cfg), C.byref(h)) == -2
    assert b"isotropic" in lib.ep_last_error()
