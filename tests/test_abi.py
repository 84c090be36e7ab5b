"""The C-ABI library loads and exports every entry point declared in include/woodem_hip.h
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "woodem_hip.h")
LIB = os.path.join(ROOT, "woodem_amd", "libwoodem_hip.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(wdm_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("wdm_create", "wdm_step", "wdm_leapfrog_run", "wdm_collider_run", "wdm_contact_loop_run",
              "wdm_get_contacts", "wdm_get_energy"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwoodem_hip.so not built")
def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwoodem_hip.so not built")
def test_params_struct_layout():
    from woodem_amd._native import WdmParams
    # doubles first (11), then 10 x 32-bit ints
    assert ctypes.sizeof(WdmParams) == 11 * 8 + 10 * 4


def test_footprint_is_monotone():
    from woodem_amd._native import Context
    if not os.path.exists(LIB):
        pytest.skip("library not built")
    a = Context.footprint(1000, 1000, 6.0)
    b = Context.footprint(2000, 2000, 6.0)
    assert 0 < a < b


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwoodem_hip.so not built")
def test_kernel_parameters_have_no_address_space():
    """A kernel parameter declared in the global address space is mangled with it (U3AS1) in
    the device pass only; the host then launches a name the code object lacks.  Every kernel
    of the built library must take plain pointers (qualified pointers only inside structs)."""
    data = open(LIB, "rb").read()
    kernels = set(m.group(1) for m in re.finditer(rb"(_Z[0-9A-Za-z_]+)\.kd\x00", data))
    ours = [k for k in kernels if re.match(rb"_Z\d+k_", k)]
    assert len(ours) > 20, "kernel symbols not found in the library's code object"
    bad = sorted(k.decode() for k in kernels if b"U3AS" in k)
    assert not bad, bad


@pytest.mark.skipif(not os.path.exists(LIB), reason="libwoodem_hip.so not built")
def test_dyndt_struct_layout():
    from woodem_amd._native import WdmDynDt
    assert ctypes.sizeof(WdmDynDt) == 6 * 8 + 5 * 8 + 4 * 4
