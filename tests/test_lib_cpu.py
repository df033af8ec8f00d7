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
