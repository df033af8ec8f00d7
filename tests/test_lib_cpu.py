"""C-ABI library checks that need no GPU: it loads, exports every symbol the
header declares, and refuses to run (loudly) when no device is visible."""
import os
import re

import pytest

from eqcycle_amd import lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_and_binding_agree():
    hdr = open(os.path.join(ROOT, "include", "eqcycle_hip.h")).read()
    declared = set(re.findall(r"^(?:int|void|const char \*)\s*(eq_\w+)\(", hdr, flags=re.M))
    assert declared == set(L.EXPORTS)


def test_library_exports_every_symbol():
    lib = L.load()
    for name in L.EXPORTS:
        assert hasattr(lib, name), name


def test_create_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from eqcycle_amd.api import EqContext, EqError
    from eqcycle_amd.configs import small
    from eqcycle_amd.params import set_parameters
    with pytest.raises(EqError):
        EqContext(set_parameters(small(2, 2)))


def test_null_context_is_rejected():
    """Every entry point that takes a context returns EQ_EINVAL for a null one
    before touching a device, as the reference's functions report misuse by
    `@error` and `return nothing` rather than crashing."""
    import ctypes as C
    lib = L.load()

    def dummy(t):
        if t in (C.c_int, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64):
            return 0
        if t in (C.c_double, C.c_float):
            return 0.0
        return None          # pointers and structs by pointer

    checked = 0
    for name in L.EXPORTS:
        if name in ("eq_create", "eq_destroy", "eq_last_error"):
            continue
        f = getattr(lib, name)
        assert f(*[dummy(t) for t in f.argtypes]) == L.EQ_EINVAL, name
        checked += 1
    assert checked == len(L.EXPORTS) - 3
    assert lib.eq_last_error(None) == b"null context"
    lib.eq_destroy(None)


def _c_struct_fields(hdr, name):
    body = re.search(r"typedef struct \{([^{}]*)\}\s*" + name + ";", hdr).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    out = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        names = re.sub(r"^(const\s+)?\w+\s*", "", decl)
        for nm in names.split(","):
            out.append(re.sub(r"\[.*\]", "", nm).strip(" *"))
    return out


def _jl_struct_fields(src, name):
    body = re.search(r"struct " + name + r"\b[^\n]*\n(.*?)\nend", src, flags=re.S).group(1)
    return re.findall(r"(\w+)::", body)


def test_julia_binding_matches_header():
    """The Julia binding calls only exported entry points, and its mirrors of
    the C structs list the same fields in the same order."""
    hdr = open(os.path.join(ROOT, "include", "eqcycle_hip.h")).read()
    jl = open(os.path.join(ROOT, "julia", "EqcycleHIP.jl")).read()
    called = set(re.findall(r"\(:(eq_\w+), LIB\)", jl))
    assert called and called <= set(L.EXPORTS), called - set(L.EXPORTS)
    for c_name, jl_name in (("eq_mesh_desc", "MeshDesc"), ("eq_params_desc", "ParamsDesc"), ("eq_csr", "EqCsr"),
                            ("eq_solver_opts", "SolverOpts"), ("eq_step_info", "StepInfo"),
                            ("eq_results", "Results")):
        assert _jl_struct_fields(jl, jl_name) == _c_struct_fields(hdr, c_name), c_name
