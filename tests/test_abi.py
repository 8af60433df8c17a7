"""CPU checks of the drop-in boundary: the built library loads and exports
every symbol declared in include/*.h; the Python, oracle and header
numberings agree.  No compute call is made (no GPU here)."""
import ctypes as C
import os
import re

import pytest

import ccake_amd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = []
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            syms += re.findall(r"^CCAKE_API\s+[\w\s\*]+?\b(ccake_\w+)\s*\(", text, flags=re.M)
    return sorted(set(syms))


def test_library_built_and_exports_every_declared_symbol():
    assert os.path.exists(ccake_amd.LIB_PATH), "build() has not produced libccake_amd.so"
    L = C.CDLL(ccake_amd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s
    # the Python front-end binds exactly the declared C-ABI
    assert sorted(ccake_amd.ABI) == syms


def test_field_numbering_matches_oracle():
    for name, v in ccake_amd.FIELDS.items():
        assert oracle.FIELDS["ORC_" + name[len("CCAKE_"):]] == v, name
    for name, v in ccake_amd.IFIELDS.items():
        assert oracle.IFIELDS["ORC_" + name[len("CCAKE_"):]] == v, name


def test_params_layout_matches_oracle():
    assert [f for f, _ in ccake_amd.PARAM_FIELDS] == [f for f, _ in oracle.PARAM_FIELDS]
    assert C.sizeof(ccake_amd.Params) == C.sizeof(oracle.OrcParams)


def test_no_device_raises_cleanly():
    """Without a HIP device the product path fails loudly (no CPU fallback)."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a HIP device is present")
    with pytest.raises(ccake_amd.CcakeError):
        ccake_amd.Device(0)
