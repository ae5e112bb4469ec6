"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("mctdhb_hip.h",):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", src):
            name = m.group(1)
            if name in ("if", "sizeof", "defined", "typedef") or name.isupper():
                continue
            # only names in declaration position (preceded by a type)
            line = src[src.rfind("\n", 0, m.start()) + 1: m.start()]
            if re.search(r"\b[A-Za-z_][A-Za-z0-9_]*\s*\**\s*$", line) and not re.search(r"\breturn\s*$", line):
                names.add(name)
    return sorted(names)


def declared_variables():
    src = open(os.path.join(ROOT, "include", "mctdhb_hip.h")).read()
    return sorted(re.findall(r"^extern\s+[A-Za-z_0-9]+\s+([A-Za-z_][A-Za-z0-9_]*)", src, flags=re.M))


def test_header_parse_finds_entry_points():
    names = declared_functions()
    for must in ("apply_hamiltonian", "set_twobody_dm", "lanczos", "propagate_coef_sil",
                 "get_multiconf_struct", "mctdhb_sil_step_dev", "mctdhb_space_create", "eig_residual",
                 "get_orbital_equation", "integration_driver", "get_builtin_pot", "get_mctdhb_from_files",
                 "mctdhb_step_graph_create", "mctdhb_last_error", "timeparam_linear_ramp"):
        assert must in names
    assert "orb_cplx_read_fmt" in declared_variables()


def test_library_exports_every_declared_symbol():
    path = os.path.join(ROOT, "1d-mctdhb_amd", "lib", "libmctdhb_hip.so")
    if not os.path.exists(path):
        pytest.fail("libmctdhb_hip.so is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    missing = [n for n in declared_functions() + declared_variables() if not hasattr(L, n)]
    assert not missing, missing
