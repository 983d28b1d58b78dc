"""C-ABI checks that need no GPU: the library loads, exports every symbol the header
declares, and fails loudly (status codes, messages) when no device is present."""
import ctypes
import os
import re

import numpy as np
import pytest

from spimagine_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spimrender.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spr_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_list():
    assert declared_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), name
    assert L.spr_abi_version() == 1


def test_params_struct_layout():
    # 16+16+6 floats, 5 floats, 4 ints
    assert ctypes.sizeof(_lib.SprParams) == (16 + 16 + 6 + 5 + 4) * 4


def test_errors_without_context():
    L = _lib.load()
    assert L.spr_set_volume(None, None, 0, 1, 1, 1) == _lib.SPR_ERR_INVALID_ARG
    assert b"NULL context" in L.spr_last_error(None)
    assert L.spr_render(None, 0, 1, 0, 4, 4, None, None, None, None, None) == _lib.SPR_ERR_INVALID_ARG
    assert L.spr_synchronize(None) == _lib.SPR_ERR_INVALID_ARG


@pytest.mark.skipif(_lib.device_count() > 0, reason="a HIP device is present")
def test_no_device_fails_loudly():
    with pytest.raises(_lib.SprError) as e:
        _lib.Context(0)
    assert e.value.code == _lib.SPR_ERR_NO_DEVICE
    from spimagine_amd.volumerender import VolumeRenderer
    with pytest.raises(_lib.SprError):
        VolumeRenderer((8, 8))


def test_dtype_codes_match_header():
    text = open(HEADER).read()
    for name, val in (("SPR_DTYPE_F32", 0), ("SPR_DTYPE_U16", 1), ("SPR_DTYPE_U8", 2),
                      ("SPR_FILTER_LINEAR", 0), ("SPR_FILTER_NEAREST", 1),
                      ("SPR_METHOD_MAX_PROJECT", 0), ("SPR_METHOD_ISO_SURFACE", 1)):
        assert re.search(r"#define %s %d\b" % (name, val), text), name
    assert _lib.DTYPE_CODES[np.dtype(np.uint16)] == 1


LEARN_HEADER = os.path.join(ROOT, "include", "spimlearn.h")


def test_learn_library_exports_every_declared_symbol():
    from spimagine_amd import learned
    text = re.sub(r"/\*.*?\*/", "", open(LEARN_HEADER).read(), flags=re.S)
    declared = sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(spl_\w+)\s*\(", text, re.M)))
    assert declared == sorted(learned.EXPORTS)
    L = learned.load()
    for name in declared:
        assert hasattr(L, name), name
    assert L.spl_abi_version() == 2
    assert L.spl_param_count() == learned.param_count() == 84996


def test_learn_errors_without_network():
    from spimagine_amd import learned
    L = learned.load()
    assert L.spl_forward(None, None, 1, 4, 4, None, None) == -1
    assert b"NULL network" in L.spl_last_error(None)
    p = learned.split_params(learned.init_params(0))
    assert not p["head.w"].any() and not p["e0.b"].any() and p["e0.w"].std() > 0
