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


def test_cpp_layer_instantiates_the_nonlinear_element(tmp_path):
    """include/gpufem/gpufem.hpp with IsotropicShell<T, Data, true>
    (include/shell/physics.h:6-17): the assembler's add_energy /
    add_residual / add_jacobian of the nonlinear element group compile
    against the C ABI (hipcc syntax check of a program that calls them)."""
    import os
    import shutil
    import subprocess

    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    if not os.path.exists(hipcc) and not shutil.which("hipcc"):
        import pytest

        pytest.skip("hipcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "nl.cpp"
    src.write_text(r'''
#include "gpufem/gpufem.hpp"
using namespace gpufem;
using T = double;
using Data = ShellIsotropicData<T, false>;
using Phys = IsotropicShell<T, Data, true>;
using Group = ShellElementGroup<T, LinearizedRotation<T>,
                                ShellQuadBasis<T, QuadLinearQuadrature<T>, 2>,
                                Phys>;
using Assembler = ElementAssembler<T, Group, VecType, BsrMat>;
static_assert(Phys::is_nonlinear, "nonlinear physics");
void use(Assembler& a, VecType<T>& u, VecType<T>& res,
         BsrMat<VecType<T>>& k, T& U) {
  a.set_variables(u);
  a.add_energy(U);
  a.add_residual(res);
  a.add_jacobian(res, k);
}
''')
    r = subprocess.run([hipcc, "-std=c++17", "-fsyntax-only",
                        "-I", os.path.join(root, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
