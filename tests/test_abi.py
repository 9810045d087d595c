"""The C-ABI library loads and exports exactly what include/nbx.h declares (no GPU needed)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nbx.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(nbx_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_the_entry_points():
    syms = declared_symbols()
    for s in ("nbx_create", "nbx_rhs", "nbx_accelerations", "nbx_vv_run", "nbx_energies",
              "nbx_rdf", "nbx_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import nbodysimulator
    from nbodysimulator import _lib
    lib = nbodysimulator.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, f"{s} has no ctypes signature"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = sorted(set(re.findall(r" T (nbx_\w+)", out)))
    assert exported == declared_symbols()


def test_library_is_built_for_gfx950_only():
    from nbodysimulator import _lib
    out = subprocess.run(["strings", _lib.LIB_PATH], capture_output=True, text=True).stdout
    targets = set(re.findall(r"amdgcn-amd-amdhsa--(gfx\w+)", out))
    assert targets == {"gfx950"}


def test_version_and_device_count_without_gpu():
    import nbodysimulator
    lib = nbodysimulator.load()
    assert lib.nbx_version().decode().startswith("nbx")
    if not os.path.exists("/dev/kfd"):
        assert nbodysimulator.device_count() == 0


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device error path")
def test_create_fails_loudly_without_device():
    import nbodysimulator
    from nbodysimulator import _lib
    with pytest.raises(_lib.NBXError) as ei:
        _lib.Context(4, _lib.SYS_POTENTIAL, [1.0] * 4)
    assert ei.value.code == _lib.NBX_ERR_DEVICE
    assert "no HIP device" in str(ei.value)


def test_argument_validation_precedes_device_use():
    import nbodysimulator
    from nbodysimulator import _lib
    lib = nbodysimulator.load()
    h = C.c_void_p()
    m = (C.c_double * 3)(1.0, 1.0, 1.0)
    assert lib.nbx_create(C.byref(h), -1, 0, m, None) == _lib.NBX_ERR_INVALID
    assert lib.nbx_create(C.byref(h), 4, 1, m, None) == _lib.NBX_ERR_INVALID  # water: 3 per mol
    assert b"3 atoms" in lib.nbx_last_error(None)
    bad = (C.c_double * 3)(1.0, -1.0, 1.0)
    assert lib.nbx_create(C.byref(h), 3, 0, bad, None) == _lib.NBX_ERR_INVALID
    assert lib.nbx_create(None, 3, 0, m, None) == _lib.NBX_ERR_INVALID
