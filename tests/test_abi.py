"""The C-ABI library loads and exports every symbol include/fwi3d.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|long|const char\*)\s+(fwi[a-z0-9_]*)\s*\(", text, re.M)))


def test_hip_library_exports_header():
    from fwi3d_amd import lib
    declared = _declared("fwi3d.h")
    assert declared, "no declarations parsed"
    L = lib.load()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    from fwi3d_amd import unet
    assert sorted(lib.EXPORTS + unet.EXPORTS) == sorted(declared)


def test_create_without_device_fails_cleanly():
    from fwi3d_amd import lib
    try:
        c = lib.Context(40, 40, 40, 8, 10.0, 1e-3, 10, 2)
    except lib.FwiError as e:
        assert e.code in (-7, -2)
        return
    c.close()
    pytest.skip("a GPU is present")


def test_create_rejects_planes_of_2_gib_and_more():
    """The step kernels address one padded x-y plane with 32-bit byte
    offsets below 2^31 (volumes of 2 GiB and more are addressed plane by
    plane): creation refuses larger planes before touching a device."""
    from fwi3d_amd import lib
    with pytest.raises(lib.FwiError) as e:
        lib.Context(32768, 16384, 9, 0, 10.0, 1e-3, 10, 2)  # 32768*16384*4 B = 2 GiB per plane
    assert e.value.code == -5  # FWI_ERR_RANGE
    with pytest.raises(lib.FwiError) as e:
        lib.Context(40, 40, 40, 8, 10.0, 1e-3, 10, 2, nshots=0)
    assert e.value.code == -1  # FWI_ERR_ARG
