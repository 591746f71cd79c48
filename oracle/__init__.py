"""CPU oracle for the linear-static MITC4 shell path -- test infrastructure.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker, never the product:
the HIP library under ``gpu_fem_amd`` neither links nor calls it.

``liboracle.so`` restates the reference host path in plain C++ (see the
citations at the top of ``element.cpp``, ``symbolic.cpp`` and
``numeric.cpp``).  ``_ref/libgfref.so`` is the reference's own
self-contained symbolic code compiled from its sources (``build_ref.sh``),
``_ref/libgfref_bdf.so`` its NASTRAN reader (``ref_bdf.cpp`` around
include/mesh/TACSMeshLoader.h); they exist only where the reference tree was
available at build time.

Every oracle function is pinned by a test in ``tests/test_oracle_*.py``
against the reference's fixtures, its compiled functions, or mathematical
invariants.  Parity status per function is recorded at the top of those
tests; functions with no such check would be marked "parity unpinned" there.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_int))


def build(force: bool = False) -> None:
    """Compile liboracle.so (and _ref when the reference tree exists)."""
    import subprocess

    if force or not os.path.exists(os.path.join(_HERE, "liboracle.so")):
        subprocess.check_call(["make", "-s", "-C", _HERE, "all"])
    if os.path.isdir("/root/reference/include"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def build_native(out_dir: str) -> str:
    """Compile the oracle with -O3 -march=native for the host it runs on
    (the CPU baseline's build; the committed build is portable) and return
    the library's path."""
    import subprocess

    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "liboracle_native.so")
    subprocess.check_call(["make", "-s", "-C", _HERE, "native",
                           f"NATIVE_OUT={out}"])
    return out


def use_library(path: str) -> None:
    """Route every oracle call through the library at ``path`` (same
    interface; used by the CPU baseline for its native build)."""
    global _LIB
    _LIB = _load(path)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _load(path)
    return _LIB


def _load(path):
    """Load an oracle build and declare its functions."""
    if True:
        L = ctypes.CDLL(path)
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_mitc4_element.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p,
                                        ctypes.POINTER(ctypes.c_double)]
        L.orc_mitc4_strain.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_mitc4_strain.restype = ctypes.c_double
        L.orc_mitc4_strain_nl.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p,
                                          _f64p, _f64p]
        L.orc_mitc4_strain_nl.restype = ctypes.c_double
        L.orc_mitc4_element_nl.argtypes = [_f64p, _f64p, _f64p,
                                           ctypes.c_void_p, _f64p,
                                           ctypes.POINTER(ctypes.c_double)]
        L.orc_assemble_tangent.argtypes = [
            ctypes.c_int, _i32p, _f64p, _f64p, _f64p, _i32p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.POINTER(ctypes.c_double)]
        L.orc_node_graph.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     _i32p, _pp, _pp]
        L.orc_rcm.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int, _i32p]
        L.orc_permute_pattern.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p,
                                          _pp, _pp]
        L.orc_iluk.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int,
                               _pp, _pp, _pp]
        L.orc_full_fill.argtypes = [ctypes.c_int, _i32p, _i32p, _pp, _pp]
        L.orc_elem_ind_map.argtypes = [ctypes.c_int, ctypes.c_int, _i32p,
                                       _i32p, _i32p, _i32p, _i32p]
        L.orc_transpose.argtypes = [ctypes.c_int, _i32p, _i32p, _pp, _pp, _pp]
        L.orc_bandwidth.argtypes = [ctypes.c_int, _i32p, _i32p]
        L.orc_min_degree.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p]
        L.orc_min_degree_stack.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p]
        L.orc_md_replay.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                    _i32p]
        L.orc_md_replay.restype = ctypes.c_longlong
        L.orc_assemble_jacobian.argtypes = [
            ctypes.c_int, _i32p, _f64p, ctypes.c_void_p, _f64p, _i32p,
            ctypes.c_void_p, _f64p, ctypes.c_void_p, ctypes.c_int]
        L.orc_assemble_residual.argtypes = [ctypes.c_int, _i32p, _f64p, _f64p,
                                            _f64p, ctypes.c_void_p, _f64p]
        L.orc_apply_bcs_mat.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p,
                                        _i32p, _i32p, _i32p, _i32p, _f64p]
        L.orc_apply_bcs_vec.argtypes = [ctypes.c_int, _i32p, ctypes.c_void_p,
                                        _f64p]
        L.orc_copy_values.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p,
                                      _i32p, _i32p, _f64p]
        L.orc_permute_vec.argtypes = [ctypes.c_int, _i32p, _f64p, _f64p]
        L.orc_bsr_ilu_f64.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p]
        L.orc_bsr_ilu_f32.argtypes = [ctypes.c_int, _i32p, _i32p, _f32p]
        L.orc_bsr_spmv_f64.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p,
                                       _f64p, _f64p]
        L.orc_bsr_trsv_f64.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p,
                                       _f64p, _f64p]
        L.orc_bsr_resid_comp_f64.argtypes = [ctypes.c_int, _i32p, _i32p,
                                             _f64p, _f64p, _f64p, _f64p]
        L.orc_bsr_trsv_f32.argtypes = [ctypes.c_int, _i32p, _i32p, _f32p,
                                       _f32p, _f32p]
        L.orc_gmres.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _i32p,
                                _i32p, _f64p, _f64p, _f64p, ctypes.c_int,
                                ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                _f64p]
        L.orc_dense_solve.argtypes = [ctypes.c_int, _f64p, _f64p, _f64p]
        L.orc_set_threads.argtypes = [ctypes.c_int]
    return L


def ref():
    """The reference's compiled symbolic functions, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libgfref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.gfref_rcm.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int,
                                ctypes.c_int, _i32p]
        L.gfref_iluk.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p,
                                 ctypes.c_int, ctypes.c_double, _pp, _pp, _pp]
        L.gfref_elem_ind_map.argtypes = [ctypes.c_int, ctypes.c_int, _i32p,
                                         ctypes.c_int, ctypes.c_int, _i32p,
                                         _i32p, _i32p, _i32p]
        L.gfref_bandwidth.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p]
        L.gfref_row_col_ptrs.argtypes = [ctypes.c_int, ctypes.c_int, _i32p,
                                         ctypes.c_int, _pp, _pp]
        L.gfref_free.argtypes = [ctypes.c_void_p]
        _REF = L
    return _REF


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _take(ptr, n, free):
    out = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].copy()
    free(ctypes.cast(ptr, ctypes.c_void_p))
    return out.astype(np.int32)


# ---------------------------------------------------------------- element
def element(xpts, vars=None, data=(70e9, 0.3, 0.005, 0.0)):
    """K (24x24), R (24), energy of one MITC4 element."""
    L = lib()
    K = np.zeros(576)
    R = np.zeros(24)
    U = ctypes.c_double()
    v = np.zeros(24) if vars is None else _f64(vars).ravel()
    L.orc_mitc4_element(_f64(xpts).ravel(), v, _f64(data), K, R,
                        ctypes.byref(U))
    return K.reshape(24, 24), R, U.value


def strain(iq, xpts, vars):
    E = np.zeros(9)
    s = lib().orc_mitc4_strain(iq, _f64(xpts).ravel(), _f64(vars).ravel(), E)
    return E, s


def strain_nl(iq, xpts, vars):
    """Nonlinear strain E (9), local displacement gradients u0x, u1x (3 x 3)
    and detXd * weight at quadrature point iq."""
    E, u0x, u1x = np.zeros(9), np.zeros(9), np.zeros(9)
    s = lib().orc_mitc4_strain_nl(iq, _f64(xpts).ravel(), _f64(vars).ravel(),
                                  E, u0x, u1x)
    return E, u0x.reshape(3, 3), u1x.reshape(3, 3), s


def element_nl(xpts, vars, data=(70e9, 0.3, 0.005, 0.0)):
    """Nonlinear element at vars: tangent K (24x24), residual R (24), energy."""
    K = np.zeros(576)
    R = np.zeros(24)
    U = ctypes.c_double()
    lib().orc_mitc4_element_nl(_f64(xpts).ravel(), _f64(vars).ravel(),
                               _f64(data), K.ctypes.data_as(ctypes.c_void_p),
                               R, ctypes.byref(U))
    return K.reshape(24, 24), R, U.value


def assemble_tangent(conn, xpts, data, elem_ind_map, perm, nnzb, vars,
                     with_k=True):
    """Nonlinear K(u) (BSR values through elem_ind_map), R(u) at
    6 perm[node] + k, and the energy."""
    conn = _i32(conn).ravel()
    ne = len(conn) // 4
    nnodes = len(np.asarray(xpts).ravel()) // 3
    kvals = np.zeros(36 * nnzb) if with_k else None
    res = np.zeros(6 * nnodes)
    U = ctypes.c_double()
    perm = None if perm is None else _i32(perm)
    lib().orc_assemble_tangent(
        ne, conn, _f64(xpts).ravel(), _f64(vars).ravel(), _f64(data).ravel(),
        _i32(elem_ind_map), _ptr(perm), _ptr(kvals), _ptr(res),
        ctypes.byref(U))
    return kvals, res, U.value


# ---------------------------------------------------------------- symbolic
def node_graph(conn, nnodes, nodes_per_elem=4):
    L = lib()
    conn = _i32(conn).ravel()
    rp, cp = ctypes.POINTER(ctypes.c_int)(), ctypes.POINTER(ctypes.c_int)()
    nnz = L.orc_node_graph(len(conn) // nodes_per_elem, nnodes,
                           nodes_per_elem, conn, ctypes.byref(rp),
                           ctypes.byref(cp))
    return _take(rp, nnodes + 1, L.orc_free), _take(cp, nnz, L.orc_free)


def rcm(rowp, cols, root=0):
    """rcm_vars[old] = new (TacsComputeRCMOrder semantics)."""
    n = len(rowp) - 1
    out = np.zeros(n, np.int32)
    lib().orc_rcm(n, _i32(rowp), _i32(cols), root, out)
    return out


def permute_pattern(rowp, cols, perm):
    L = lib()
    n = len(rowp) - 1
    rp, cp = ctypes.POINTER(ctypes.c_int)(), ctypes.POINTER(ctypes.c_int)()
    nnz = L.orc_permute_pattern(n, _i32(rowp), _i32(cols), _i32(perm),
                                ctypes.byref(rp), ctypes.byref(cp))
    return _take(rp, n + 1, L.orc_free), _take(cp, nnz, L.orc_free)


def iluk(rowp, cols, k):
    L = lib()
    n = len(rowp) - 1
    rp, cp, lv = (ctypes.POINTER(ctypes.c_int)() for _ in range(3))
    nnz = L.orc_iluk(n, _i32(rowp), _i32(cols), k, ctypes.byref(rp),
                     ctypes.byref(cp), ctypes.byref(lv))
    return (_take(rp, n + 1, L.orc_free), _take(cp, nnz, L.orc_free),
            _take(lv, nnz, L.orc_free))


def full_fill(rowp, cols):
    L = lib()
    n = len(rowp) - 1
    rp, cp = ctypes.POINTER(ctypes.c_int)(), ctypes.POINTER(ctypes.c_int)()
    nnz = L.orc_full_fill(n, _i32(rowp), _i32(cols), ctypes.byref(rp),
                          ctypes.byref(cp))
    return _take(rp, n + 1, L.orc_free), _take(cp, nnz, L.orc_free)


def elem_ind_map(conn, rowp, cols, perm, nodes_per_elem=4):
    conn = _i32(conn).ravel()
    ne = len(conn) // nodes_per_elem
    out = np.zeros(ne * nodes_per_elem ** 2, np.int32)
    lib().orc_elem_ind_map(ne, nodes_per_elem, conn, _i32(rowp), _i32(cols),
                           _i32(perm), out)
    return out


def transpose(rowp, cols):
    L = lib()
    n = len(rowp) - 1
    rp, cp, mp = (ctypes.POINTER(ctypes.c_int)() for _ in range(3))
    nnz = L.orc_transpose(n, _i32(rowp), _i32(cols), ctypes.byref(rp),
                          ctypes.byref(cp), ctypes.byref(mp))
    return (_take(rp, n + 1, L.orc_free), _take(cp, nnz, L.orc_free),
            _take(mp, nnz, L.orc_free))


def bandwidth(rowp, cols):
    return lib().orc_bandwidth(len(rowp) - 1, _i32(rowp), _i32(cols))


def min_degree(rowp, cols):
    n = len(rowp) - 1
    out = np.zeros(n, np.int32)
    lib().orc_min_degree(n, _i32(rowp), _i32(cols), out)
    return out


def min_degree_stack(rowp, cols):
    """Exact minimum degree, ties to the most recently updated node."""
    n = len(rowp) - 1
    out = np.zeros(n, np.int32)
    lib().orc_min_degree_stack(n, _i32(rowp), _i32(cols), out)
    return out


def md_replay(rowp, cols, perm):
    """Replays an ordering on the explicit elimination graph: (exact degree
    of each pivot, least exact degree at that step, nnz(L) blocks)."""
    n = len(rowp) - 1
    dp = np.zeros(n, np.int32)
    dm = np.zeros(n, np.int32)
    nnz = lib().orc_md_replay(n, _i32(rowp), _i32(cols), _i32(perm), dp, dm)
    return dp, dm, int(nnz)


# ---------------------------------------------------------------- numeric
def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def assemble_jacobian(conn, xpts, data, elem_ind_map, perm, nnzb, vars=None,
                      with_res=True):
    conn = _i32(conn).ravel()
    ne = len(conn) // 4
    nnodes = len(np.asarray(xpts).ravel()) // 3
    kvals = np.zeros(36 * nnzb)
    res = np.zeros(6 * nnodes) if with_res else None
    v = None if vars is None else _f64(vars).ravel()
    perm = _i32(perm)
    lib().orc_assemble_jacobian(ne, conn, _f64(xpts).ravel(), _ptr(v),
                                _f64(data).ravel(), _i32(elem_ind_map),
                                _ptr(perm), kvals, _ptr(res), 1)
    return kvals, res


def assemble_residual(conn, xpts, vars, data, perm=None):
    conn = _i32(conn).ravel()
    nnodes = len(np.asarray(xpts).ravel()) // 3
    res = np.zeros(6 * nnodes)
    p = None if perm is None else _i32(perm)
    lib().orc_assemble_residual(len(conn) // 4, conn, _f64(xpts).ravel(),
                                _f64(vars).ravel(), _f64(data).ravel(),
                                _ptr(p), res)
    return res


def apply_bcs_mat(bcs, rowp, cols, trowp, tcols, tmap, perm, vals):
    vals = _f64(vals).copy()
    bcs = _i32(bcs)
    lib().orc_apply_bcs_mat(len(bcs), bcs, _i32(rowp), _i32(cols),
                            _i32(trowp), _i32(tcols), _i32(tmap), _i32(perm),
                            vals)
    return vals


def apply_bcs_vec(bcs, vec, perm=None):
    vec = _f64(vec).copy()
    bcs = _i32(bcs)
    p = None if perm is None else _i32(perm)
    lib().orc_apply_bcs_vec(len(bcs), bcs, _ptr(p), vec)
    return vec


def copy_values(rowp, cols, vals, trowp, tcols):
    tv = np.zeros(36 * int(trowp[-1]))
    lib().orc_copy_values(len(rowp) - 1, _i32(rowp), _i32(cols), _f64(vals),
                          _i32(trowp), _i32(tcols), tv)
    return tv


def permute_vec(perm, v):
    out = np.zeros_like(_f64(v))
    lib().orc_permute_vec(len(perm), _i32(perm), _f64(v), out)
    return out


def bsr_ilu(rowp, cols, vals, dtype=np.float64, rows=None):
    """Block ILU in place on a copy; ``rows``: factor only rows [0, rows)
    (FP64; a timing sample, exact for those rows)."""
    v = np.ascontiguousarray(vals, dtype=dtype).copy()
    n = len(rowp) - 1
    if rows is not None:
        L = lib()
        L.orc_bsr_ilu_rows_f64.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p,
                                           ctypes.c_int]
        return v, L.orc_bsr_ilu_rows_f64(n, _i32(rowp), _i32(cols), v,
                                         int(rows))
    if dtype == np.float64:
        st = lib().orc_bsr_ilu_f64(n, _i32(rowp), _i32(cols), v)
    else:
        st = lib().orc_bsr_ilu_f32(n, _i32(rowp), _i32(cols), v)
    return v, st


def bsr_spmv(rowp, cols, vals, x):
    y = np.zeros(6 * (len(rowp) - 1))
    lib().orc_bsr_spmv_f64(len(rowp) - 1, _i32(rowp), _i32(cols), _f64(vals),
                           _f64(x), y)
    return y


def bsr_resid_comp(rowp, cols, vals, x, b):
    """b - A x with compensated (twice-FP64) accumulation."""
    r = np.zeros(6 * (len(rowp) - 1))
    lib().orc_bsr_resid_comp_f64(len(rowp) - 1, _i32(rowp), _i32(cols),
                                 _f64(vals), _f64(x), _f64(b), r)
    return r


def bsr_trsv(rowp, cols, lu, b, dtype=np.float64):
    n = len(rowp) - 1
    if dtype == np.float64:
        x = np.zeros(6 * n)
        lib().orc_bsr_trsv_f64(n, _i32(rowp), _i32(cols), _f64(lu), _f64(b), x)
    else:
        x = np.zeros(6 * n, np.float32)
        lib().orc_bsr_trsv_f32(n, _i32(rowp), _i32(cols),
                               np.ascontiguousarray(lu, np.float32),
                               np.ascontiguousarray(b, np.float32), x)
    return x


def gmres(krowp, kcols, kvals, mrowp, mcols, mlu, b, restart=50,
          max_iters=2000, rtol=1e-8, mixed=False, x0=None):
    n = len(krowp) - 1
    x = np.zeros(6 * n) if x0 is None else _f64(x0).copy()
    stats = np.zeros(4)
    st = lib().orc_gmres(n, _i32(krowp), _i32(kcols), _f64(kvals),
                         _i32(mrowp), _i32(mcols), _f64(mlu), _f64(b), x,
                         restart, max_iters, rtol, int(mixed), stats)
    return x, st, {"iters": int(stats[0]), "rel_resid": float(stats[1]),
                   "outer": int(stats[2])}


def dense_solve(A, b):
    A = _f64(A)
    x = np.zeros(A.shape[0])
    st = lib().orc_dense_solve(A.shape[0], A.ravel(), _f64(b), x)
    if st != 0:
        raise np.linalg.LinAlgError("singular")
    return x


# ------------------------------------------------- reference-compiled (_ref)
def ref_rcm(rowp, cols, root=0, iters=1):
    L = ref()
    n = len(rowp) - 1
    out = np.zeros(n, np.int32)
    L.gfref_rcm(n, _i32(rowp), _i32(cols), root, iters, out)
    return out


def ref_iluk(rowp, cols, k, fill=10.0):
    L = ref()
    n = len(rowp) - 1
    rp, cp, lv = (ctypes.POINTER(ctypes.c_int)() for _ in range(3))
    nnz = L.gfref_iluk(n, int(rowp[-1]), _i32(rowp), _i32(cols), k, fill,
                       ctypes.byref(rp), ctypes.byref(cp), ctypes.byref(lv))
    return (_take(rp, n + 1, L.gfref_free), _take(cp, nnz, L.gfref_free),
            _take(lv, nnz, L.gfref_free))


def ref_elem_ind_map(conn, nnodes, rowp, cols, perm, nodes_per_elem=4):
    L = ref()
    conn = _i32(conn).ravel()
    ne = len(conn) // nodes_per_elem
    out = np.zeros(ne * nodes_per_elem ** 2, np.int32)
    L.gfref_elem_ind_map(ne, nnodes, conn, nodes_per_elem, int(rowp[-1]),
                         _i32(rowp).copy(), _i32(cols).copy(),
                         _i32(perm).copy(), out)
    return out


def ref_bandwidth(rowp, cols):
    return ref().gfref_bandwidth(len(rowp) - 1, int(rowp[-1]),
                                 _i32(rowp).copy(), _i32(cols).copy())


def ref_node_graph(conn, nnodes, nodes_per_elem=4):
    L = ref()
    conn = _i32(conn).ravel()
    rp, cp = ctypes.POINTER(ctypes.c_int)(), ctypes.POINTER(ctypes.c_int)()
    nnz = L.gfref_row_col_ptrs(len(conn) // nodes_per_elem, nnodes, conn,
                               nodes_per_elem, ctypes.byref(rp),
                               ctypes.byref(cp))
    return _take(rp, nnodes + 1, L.gfref_free), _take(cp, nnz, L.gfref_free)


_REF_BDF = None


def ref_bdf_lib():
    """The reference's NASTRAN reader (``_ref/libgfref_bdf.so``, built from
    include/mesh/TACSMeshLoader.h by ``build_ref.sh``), or None."""
    global _REF_BDF
    if _REF_BDF is None:
        path = os.path.join(_HERE, "_ref", "libgfref_bdf.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.gfref_bdf_scan.restype = ctypes.c_void_p
        L.gfref_bdf_scan.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.gfref_bdf_sizes.argtypes = [ctypes.c_void_p, _i32p]
        L.gfref_bdf_get.argtypes = [ctypes.c_void_p, _i32p, ctypes.c_int,
                                    _i32p, _f64p]
        L.gfref_bdf_free.argtypes = [ctypes.c_void_p]
        _REF_BDF = L
    return _REF_BDF


def ref_bdf(path, vars_per_node=6, nodes_per_elem=4):
    """Reference reader on a file: None if it reports failure, else
    (conn [nelems, nodes_per_elem], global BC DOFs, xpts [nnodes, 3]) from
    TACSMeshLoader::getAssemblerCreatorData."""
    L = ref_bdf_lib()
    h = L.gfref_bdf_scan(os.fsencode(path), vars_per_node)
    try:
        sz = np.zeros(4, np.int32)
        L.gfref_bdf_sizes(h, sz)
        fail, nn, ne, nb = (int(v) for v in sz)
        if fail:
            return None
        conn = np.zeros(nodes_per_elem * ne, np.int32)
        bcs = np.zeros(nb, np.int32)
        x = np.zeros(3 * nn)
        L.gfref_bdf_get(h, conn, len(conn), bcs, x)
        return conn.reshape(ne, nodes_per_elem), bcs, x.reshape(nn, 3)
    finally:
        L.gfref_bdf_free(h)
