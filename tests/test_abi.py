"""The C-ABI library loads without a GPU and exports every entry point of include/sph_hip.h."""
import ctypes
import os

import pytest

from sphfluidsimulation_amd import _lib


def test_library_built_and_loads():
    assert os.path.exists(_lib.LIB_PATH), "run __graft_entry__.build() first"
    lib = _lib.load()
    assert lib.sph_system_implementation_name().decode() == "HIP"


def test_exports_every_header_symbol():
    lib = _lib.load()
    declared = _lib.header_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # and the ctypes table covers the header exactly
    assert sorted(_lib.SIGNATURES) == declared


def test_no_device_is_an_error_not_a_fallback():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.sph_create(ctypes.byref(h), 0)
    if rc == _lib.SPH_OK:  # a HIP device is present
        lib.sph_destroy(h)
        pytest.skip("a HIP device is present")
    assert rc == _lib.SPH_E_NO_DEVICE


def test_null_handle_is_rejected():
    lib = _lib.load()
    assert lib.sph_update(None, 1e-3, 1) == _lib.SPH_E_INVALID
    assert lib.sph_synchronize(None) == _lib.SPH_E_INVALID
    assert lib.sph_read(None, None, None, None) == _lib.SPH_E_INVALID
    assert lib.sph_last_error(None) == b"invalid handle"


def test_header_constants_match_the_host():
    """Every #define SPH_* of include/sph_hip.h that the Python host names has the header's value."""
    import re

    text = open(_lib.HEADER_PATH).read()
    defines = {}
    for name, expr in re.findall(r"#define\s+(SPH_[A-Z0-9_]+)[ \t]+([^/\n]+)", text):
        expr = re.sub(r"(?<=[0-9a-fA-F])[uU]\b", "", expr.strip())
        if re.fullmatch(r"[0-9xXa-fA-F()<\s-]+", expr):
            defines[name] = eval(expr, {"__builtins__": {}})  # integer literals, shifts and parentheses only
    assert defines["SPH_PROFILE_ROTATE"] == 0x10 and defines["SPH_NOALLOC_MAX_PARTICLES"] == 1 << 20
    checked = 0
    for name, value in defines.items():
        if hasattr(_lib, name):
            assert getattr(_lib, name) == value, name
            checked += 1
    assert checked >= 15
