"""The C-ABI library loads and exports every entry point that
include/gpufem_amd.h declares (no device calls)."""
import ctypes

import gpu_fem_amd as G


def test_library_exports_every_declared_symbol():
    L = G.lib()
    names = G.exported_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.gf_version() >= 1


def test_python_binding_covers_header():
    assert set(G.exported_symbols()) == set(G._SIGS)


def test_invalid_arguments_are_rejected_without_device():
    L = G.lib()
    out = ctypes.c_void_p()
    assert L.gf_context_create(-1, ctypes.byref(out)) != G.GF_OK
    assert L.gf_rcm_order(0, None, None, 0, 1, None) == -1
    assert L.gf_bandwidth(0, None, None) == -1
    o = G.default_options()
    assert (o.reorder, o.restart, o.precision) == (1, 50, 0)
