"""The C ABI library: it loads without a GPU and exports every entry point include/relearn_bh.h declares.

No compute call is made here (there is no GPU in the CPU suite); the GPU tests call through the same
ABI.  Also checks that the ctypes structures match the header's struct layouts field by field.
"""
import ctypes
import os
import re

import pytest

from relearn_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "relearn_bh.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(rl_[a-z0-9_]+)\s*\(", text, re.M)))


def header_struct_fields(name):
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    m = None
    for cand in re.finditer(r"typedef\s+struct\s*(?:[A-Za-z_][A-Za-z0-9_]*)?\s*\{([^{}]*)\}\s*([A-Za-z_][A-Za-z0-9_]*)\s*;", text):
        if cand.group(2) == name:
            m = cand
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        fm = re.match(r"((?:const\s+)?[A-Za-z_][A-Za-z0-9_]*)\s+(.+)$", decl)
        assert fm, decl
        for d in fm.group(2).split(","):
            nm = re.match(r"\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*(\[\s*\d+\s*\])?\s*$", d)
            assert nm, decl
            fields.append(nm.group(1))
    return fields


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("librelearn_bh.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(_lib.LIB_PATH)


def test_header_declares_the_entry_points():
    names = declared_functions()
    assert "rl_bh_create" in names and "rl_bh_find_target_neurons" in names and "rl_neurons_step" in names
    assert len(names) >= 20


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_the_header():
    # the ctypes binding names exactly the header's entry points
    assert sorted(_lib.EXPORTED) == declared_functions()


@pytest.mark.parametrize("cname,pyname", [("rl_bh_config", "Config"), ("rl_bh_step_stats", "StepStats"),
                                          ("rl_step_params", "StepParams")])
def test_struct_layouts_match_header(cname, pyname):
    py = getattr(_lib, pyname)
    assert [f[0] for f in py._fields_] == header_struct_fields(cname)


def test_create_rejects_bad_config_without_gpu(lib):
    # argument validation happens before any device call
    lib.rl_bh_create.restype = ctypes.c_int
    lib.rl_bh_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    out = ctypes.c_void_p()
    assert lib.rl_bh_create(None, ctypes.byref(out)) != 0
    lib.rl_bh_last_error.restype = ctypes.c_char_p
    lib.rl_bh_last_error.argtypes = [ctypes.c_void_p]
    cfg = _lib.Config()
    cfg.box_min[:] = (0.0, 0.0, 0.0)
    cfg.box_max[:] = (1.0, 1.0, 1.0)
    cfg.theta, cfg.sigma, cfg.mu = 0.3, 750.0, 0.0
    cfg.precision = _lib.RL_PREC_FP32_MIXED
    cfg.flags = 16
    assert lib.rl_bh_create(ctypes.byref(cfg), ctypes.byref(out)) == -1
    assert b"flag" in lib.rl_bh_last_error(None)
    cfg.flags = _lib.RL_FLAG_COUNT_WALK
    cfg.theta = 0.7
    assert lib.rl_bh_create(ctypes.byref(cfg), ctypes.byref(out)) == -1
    assert b"theta" in lib.rl_bh_last_error(None)
