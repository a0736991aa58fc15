"""The C-ABI library loads and exports every entry point include/wave_abi.h declares (no GPU needed)."""
import ctypes
import os
import re

import wavesim._abi as abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "wave_abi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(wv_\w+)\s*\(", text, re.M)))


def test_header_matches_binding_list():
    assert _declared() == sorted(abi.EXPORTS)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name


def test_last_error_is_callable_without_gpu():
    lib = abi.load()
    assert isinstance(lib.wv_last_error(), bytes)


def test_factory_rejects_unknown_types_without_gpu():
    """ForwardSolverFactory.cpp:4-66 throws on unknown dimension/type; the ABI returns WV_EINVAL."""
    lib = abi.load()
    h = ctypes.c_void_p()
    assert lib.wv_create(b"4D", b"elastic", ctypes.byref(h)) == -1
    assert b"dimension" in lib.wv_last_error()
    assert lib.wv_create(b"2D", b"plasma", ctypes.byref(h)) == -1
    # factory types this build does not provide
    # (the reference factory returns an empty pointer for 3D tmem / sh / viscosh / viscotmem)
    for d, t in ((b"3D", b"tmem"), (b"3D", b"sh"), (b"3D", b"viscosh"), (b"3D", b"viscotmem")):
        assert lib.wv_create(d, t, ctypes.byref(h)) == -1, (d, t)
        assert b"not provided" in lib.wv_last_error()
    # and the ones it does (creation touches no GPU)
    for d, t in ((b"2D", b"viscoelastic"), (b"2D", b"viscosh"), (b"2D", b"sh"), (b"2D", b"emem"), (b"3D", b"acoustic"),
                 (b"3D", b"viscoelastic"), (b"2D", b"viscotmem"), (b"2D", b"viscoemem")):
        assert lib.wv_create(d, t, ctypes.byref(h)) == 0, (d, t)
        lib.wv_destroy(h)
