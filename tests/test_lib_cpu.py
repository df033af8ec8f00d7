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

    skip = ("eq_create", "eq_destroy", "eq_last_error", "eq_rs_free")
    checked = 0
    for name in L.EXPORTS:
        if name in skip:
            continue
        f = getattr(lib, name)
        assert f(*[dummy(t) for t in f.argtypes]) == L.EQ_EINVAL, name
        checked += 1
    assert checked == len(L.EXPORTS) - len(skip)
    assert lib.eq_last_error(None) == b"null context"
    lib.eq_destroy(None)
    lib.eq_rs_free(None)


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
                            ("eq_results", "Results"), ("eq_fault_params", "FaultParams")):
        assert _jl_struct_fields(jl, jl_name) == _c_struct_fields(hdr, c_name), c_name
    # every positional construction of a struct mirror passes one value per field
    for jl_name in ("SolverOpts",):
        nf = len(_jl_struct_fields(jl, jl_name))
        calls = re.findall(r"\b" + jl_name + r"\(([^()]*)\)", jl)
        assert calls, jl_name
        for args in calls:
            assert len([a for a in args.split(",") if a.strip()]) == nf, (jl_name, args)


# The reference routines' signatures (name, positional arguments, defaults):
# src/NRsearch_serial.jl:6,47, src/otherFunctions_serial.jl:5,35,62,
# src/dtevol.jl:5, src/main.jl:67.
REFERENCE_SIGNATURES = {
    "FBC!": ("IDstate", "P", "NFBC", "FltNglob", "psi1", "Vf1", "tau1", "psi2", "Vf2", "tau2", "psi", "Vf",
             "FltVfree", "dt"),
    "NRsearch!": ("fo", "Vo", "cca", "ccb", "Seff", "tau", "tauo", "psi", "FltZ", "FltVfree"),
    "IDS!": ("xLf", "Vo", "psi", "dt", "Vf", "cnd", "IDstate = 2"),
    "IDS2!": ("xLf", "Vo", "psi", "psi1", "dt", "Vf", "Vf1", "IDstate = 2"),
    "slrFunc!": ("P", "NFBC", "FltNglob", "psi", "psi1", "Vf", "Vf1", "IDstate", "tau1", "dt"),
    "dtevol!": ("dt", "dtmin", "XiLf", "FaultNglob", "NFBC", "Vf", "isolver"),
    "main": ("P",),
}


def test_julia_binding_has_the_reference_signatures():
    """Each routine has a method with the reference's positional arguments
    (same names and order, same defaults), so a call site needs no edit."""
    jl = open(os.path.join(ROOT, "julia", "EqcycleHIP.jl")).read()
    for name, args in REFERENCE_SIGNATURES.items():
        pat = r"^" + re.escape(name) + r"\(([^)]*)\)\s*="
        found = [tuple(a.strip() for a in m.split(",")) for m in re.findall(pat, jl, flags=re.M)]
        assert args in found, (name, found)


def test_python_mirror_has_the_reference_signatures():
    import inspect

    from eqcycle_amd import api
    for name, args in REFERENCE_SIGNATURES.items():
        f = getattr(api, name.rstrip("!"))
        got = []
        for p in inspect.signature(f).parameters.values():
            got.append(p.name if p.default is inspect.Parameter.empty else f"{p.name} = {p.default}")
        if name == "main":
            got = got[:1]     # main(P) plus keyword-only extras of the device loop
        assert tuple(got) == args, (name, got)

