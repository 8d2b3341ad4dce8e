"""The C-ABI library loads and exports every declared entry point; host logic
that needs no GPU (parameter defaults, env mapping, error paths)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hx.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(hx_[a-z0-9_]+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("hx_build", "hx_rebuild", "hx_refit", "hx_trace", "hx_trace_sorted", "hx_bvh_get_view",
                 "hx_sort_pairs_u64", "hx_morton_codes", "hx_context_create"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from hippie_amd import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes signature table covers them all
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_status_strings_and_defaults():
    from hippie_amd import _lib
    L = _lib.lib()
    assert L.hx_status_string(0) == b"ok"
    assert b"unsupported" in L.hx_status_string(_lib.HX_ERR_UNSUPPORTED)
    p = _lib.HxBuildParams()
    assert L.hx_build_params_default(ctypes.byref(p)) == 0
    # environment/AppEnvironment.cpp:52-75 defaults
    assert (p.method, p.morton_bits, p.max_leaf_size, p.adaptive_leaf_size) == (0, 60, 8, 1)
    assert (p.ct, p.ci) == (3.0, 2.0)
    assert (p.ploc_radius, p.ploc_max_radius, p.atr_iterations, p.hlbvh_sah_bits) == (1, 512, 2, 15)
    assert abs(p.presplitter_beta - 0.1) < 1e-7 and p.layout == _lib.HX_LAYOUT_BIN


def test_null_arguments_are_rejected_without_a_device():
    from hippie_amd import _lib
    L = _lib.lib()
    assert L.hx_build_params_default(None) == _lib.HX_ERR_INVALID_ARG
    assert L.hx_context_create(0, None) == _lib.HX_ERR_INVALID_ARG
    assert L.hx_bvh_get_view(None, None) == _lib.HX_ERR_INVALID_ARG
    assert L.hx_trace(None, None, None, 0, 0, None, None) == _lib.HX_ERR_INVALID_ARG
    assert b"NULL" in L.hx_last_error()


def test_env_mapping():
    pytest.importorskip("torch")
    from hippie_amd.core import BuildParams
    p = BuildParams.from_env({"Bvh.method": "lbvh", "Bvh.lbvhMortonCodeBits": "30", "Bvh.maxLeafSize": "4",
                              "Bvh.adaptiveLeafSize": "0", "Bvh.ct": "1.5"})
    assert (p.morton_bits, p.max_leaf_size, p.adaptive_leaf_size) == (30, 4, 0)
    assert p.ct == 1.5
    p = BuildParams.from_env({"Bvh.method": "ploc", "Bvh.plocMortonCodeBits": "45", "Bvh.plocRadius": "25"})
    assert (p.method, p.morton_bits, p.ploc_radius) == (2, 45, 25)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr
