"""gpu_fem_amd: MI355X (gfx950) linear-static MITC4 shell path.

The product is the C-ABI shared library ``libgpufem_amd.so`` (declared in
``include/gpufem_amd.h``) and the C++ host classes in ``include/gpufem/``.
This module loads the library with ctypes for Python callers (tests, the
benchmark); device buffers are passed as raw pointers (torch tensors are
used only as device allocations).  There is no CPU fallback: every device
entry point requires the HIP library, and loading fails loudly when it is
missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import meshes  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpufem_amd.so")
_LIB = None

GF_OK = 0
STATUS = {0: "GF_OK", -1: "GF_ERR_INVALID", -2: "GF_ERR_NOMEM",
          -3: "GF_ERR_DEVICE", -4: "GF_ERR_ZERO_PIVOT", -5: "GF_ERR_BREAKDOWN",
          -6: "GF_ERR_NO_CONVERGENCE", -7: "GF_ERR_OVER_BUDGET",
          -8: "GF_ERR_TIMEOUT"}


class GfError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("reorder", ctypes.c_int), ("lev_fill", ctypes.c_int),
                ("restart", ctypes.c_int), ("max_iters", ctypes.c_int),
                ("rtol", ctypes.c_double), ("precision", ctypes.c_int),
                ("stall_cycles", ctypes.c_int)]


class NewtonInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int),
                ("newton_iterations", ctypes.c_int),
                ("krylov_iterations", ctypes.c_int),
                ("rel_resid", ctypes.c_double), ("t_total", ctypes.c_double)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int),
                ("outer", ctypes.c_int), ("rel_resid", ctypes.c_double),
                ("t_assemble", ctypes.c_double), ("t_factor", ctypes.c_double),
                ("t_solve", ctypes.c_double)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_PI = ctypes.POINTER(ctypes.c_int)
_PPI = ctypes.POINTER(_PI)

# name -> argtypes (restype int unless listed in _RESTYPES)
_SIGS = {
    "gf_version": [],
    "gf_context_create": [_I, ctypes.POINTER(_P)],
    "gf_context_destroy": [_P],
    "gf_context_synchronize": [_P],
    "gf_context_last_error": [_P],
    "gf_context_stream": [_P],
    "gf_free": [_P],
    "gf_node_graph": [_I, _I, _I, _P, _PPI, _PPI, _PI],
    "gf_rcm_order": [_I, _P, _P, _I, _I, _P],
    "gf_amd_order": [_I, _P, _P, _P],
    "gf_pattern_permute": [_I, _P, _P, _P, _PPI, _PPI],
    "gf_iluk_symbolic": [_I, _P, _P, _I, _PPI, _PPI, _PI],
    "gf_elem_ind_map": [_I, _I, _P, _P, _P, _P, _P],
    "gf_transpose_map": [_I, _P, _P, _PPI, _PPI, _PPI],
    "gf_bandwidth": [_I, _P, _P],
    "gf_assemble_jacobian_mitc4_f64": [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P,
                                       _P],
    "gf_assemble_residual_mitc4_f64": [_P, _I, _I, _P, _P, _P, _P, _P, _P],
    "gf_asm_plan_create": [_P, _I, _I, _P, _P, _I, ctypes.POINTER(_P)],
    "gf_asm_plan_destroy": [_P],
    "gf_asm_plan_colors": [_P],
    "gf_assemble_jacobian_colored_f64": [_P, _P, _P, _P, _P, _P, _P, _P, _P,
                                         _P],
    "gf_energy_mitc4_f64": [_P, _I, _P, _P, _P, _P, _P],
    "gf_assemble_tangent_mitc4_f64": [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P,
                                      _P, _P],
    "gf_assemble_tangent_colored_f64": [_P, _P, _P, _P, _P, _P, _P, _P, _P,
                                        _P],
    "gf_vec_axpby_f64": [_P, ctypes.c_size_t, ctypes.c_double, _P,
                         ctypes.c_double, _P],
    "gf_bc_apply_mat": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "gf_bc_apply_vec": [_P, _I, _P, _P, _P],
    "gf_bsr_copy_to_pattern": [_P, _I, _P, _P, _P, _P, _P, _P],
    "gf_permute_vec": [_P, _I, _P, _P, _P],
    "gf_bsr_spmv_f64": [_P, _I, _P, _P, _P, _P, _P],
    "gf_bsr_spmv_f32": [_P, _I, _P, _P, _P, _P, _P],
    "gf_convert_f64_f32": [_P, ctypes.c_size_t, _P, _P],
    "gf_bsr_resid_comp_f64": [_P, _I, _P, _P, _P, _P, _P, _P],
    "gf_bsr_resid_f64": [_P, _I, _P, _P, _P, _P, _P,
                         ctypes.POINTER(ctypes.c_double),
                         ctypes.POINTER(ctypes.c_double)],
    "gf_bsr_resid_f64_noalloc": [_P, _I, _P, _P, _P, _P, _P, _P],
    "gf_ilu_create": [_P, _I, _P, _P, ctypes.POINTER(_P)],
    "gf_ilu_destroy": [_P],
    "gf_ilu_info": [_P, _PI, _PI, _PI],
    "gf_ilu_device_rowp": [_P],
    "gf_ilu_device_cols": [_P],
    "gf_ilu_check": [_P],
    "gf_bsr_ilu_factor_f64": [_P, _P, _PI],
    "gf_bsr_ilu_factor_f32": [_P, _P, _PI],
    "gf_bsr_ilu_apply_f64": [_P, _P, _P, _P, _P],
    "gf_bsr_ilu_apply_f32": [_P, _P, _P, _P, _P],
    "gf_gmres_solve": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I,
                       ctypes.c_double, _I, _PI, ctypes.POINTER(ctypes.c_double)],
    "gf_options_default": [ctypes.POINTER(Options)],
    "gf_problem_create": [_P, _I, _I, _P, _P, _P, _I, _P,
                          ctypes.POINTER(Options), ctypes.POINTER(_P)],
    "gf_problem_destroy": [_P],
    "gf_problem_assemble": [_P],
    "gf_problem_factor": [_P],
    "gf_problem_solve": [_P, _P, _P, _I, ctypes.POINTER(SolveInfo)],
    "gf_problem_sizes": [_P, _PI, _PI, _PI],
    "gf_problem_assemble_tangent": [_P, _P, _P, _I],
    "gf_problem_newton_solve": [_P, _P, _P, _I, _I, ctypes.c_double, _P,
                                ctypes.POINTER(NewtonInfo)],
    "gf_problem_device_arrays": [_P] + [ctypes.POINTER(_P)] * 7,
    "gf_problem_host_perm": [_P, _P],
    "gf_problem_ilu": [_P],
    "gf_server_create": [_I, ctypes.c_double, _I, ctypes.POINTER(_P)],
    "gf_server_submit": [_P, ctypes.c_longlong, _I, _I, _P, _P, _P, _I, _P,
                         _P, ctypes.POINTER(Options)],
    "gf_server_poll": [_P, ctypes.POINTER(ctypes.c_longlong), _P,
                       ctypes.c_size_t, ctypes.POINTER(SolveInfo),
                       ctypes.POINTER(ctypes.c_double)],
    "gf_server_stats": [_P, _P],
    "gf_server_drain": [_P],
    "gf_server_destroy": [_P],
    "gf_bdf_read": [ctypes.c_char_p, ctypes.POINTER(_P)],
    "gf_bdf_error": [],
    "gf_bdf_destroy": [_P],
    "gf_bdf_sizes": [_P, _PI, _PI, _PI, _PI, _PI, _PI],
    "gf_bdf_connectivity": [_P, _P, _P, _P, _P],
    "gf_bdf_bcs": [_P, _P, _P, _P, _P],
    "gf_bdf_assembler_data": [_P, _I, _P, _P, _P],
    "gf_bdf_component_descript": [_P, _I],
    "gf_bdf_element_descript": [_P, _I],
    "gf_write_vtk": [ctypes.c_char_p, _I, _P, _I, _P, _P],
}
_RESTYPES = {"gf_context_last_error": ctypes.c_char_p,
             "gf_context_stream": _P, "gf_free": None,
             "gf_options_default": None, "gf_ilu_device_rowp": _P,
             "gf_problem_ilu": _P,
             "gf_ilu_device_cols": _P, "gf_bdf_error": ctypes.c_char_p,
             "gf_bdf_component_descript": ctypes.c_char_p,
             "gf_bdf_element_descript": ctypes.c_char_p}


def exported_symbols():
    """Names declared in include/gpufem_amd.h (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "gpufem_amd.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(gf_[a-z0-9_]+)\s*\(", txt)))


def build(force=False):
    from .build import build as _b
    return _b(force=force)


def lib():
    """Load libgpufem_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: run __graft_entry__.build() or "
                "python gpu_fem_amd/build.py (there is no CPU fallback)")
        # torch ships its own HIP and HSA runtimes; loading this library
        # first would map the system ones as well, and two HSA runtimes in
        # one process conflict over the device.  With torch loaded first the
        # library binds to torch's runtime (same sonames).
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def check(code, ctx=None):
    if code != GF_OK:
        msg = ""
        if ctx is not None:
            msg = lib().gf_context_last_error(ctx).decode()
        raise GfError(code, msg)
    return code


def ptr(a):
    """Raw pointer of a numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(_P)
    return _P(a.data_ptr())


def _take_host(p, n):
    out = np.ctypeslib.as_array(p, shape=(max(n, 1),))[:n].copy()
    lib().gf_free(ctypes.cast(p, _P))
    return out.astype(np.int32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# --------------------------------------------------------- host symbolic
def node_graph(conn, nnodes, nodes_per_elem=4):
    conn = _i32(conn).ravel()
    rp, cp, nnz = _PI(), _PI(), ctypes.c_int()
    check(lib().gf_node_graph(len(conn) // nodes_per_elem, nnodes,
                              nodes_per_elem, ptr(conn), ctypes.byref(rp),
                              ctypes.byref(cp), ctypes.byref(nnz)))
    return _take_host(rp, nnodes + 1), _take_host(cp, nnz.value)


def rcm_order(rowp, cols, root=0, iters=1):
    rowp, cols = _i32(rowp), _i32(cols)
    out = np.zeros(len(rowp) - 1, np.int32)
    check(lib().gf_rcm_order(len(rowp) - 1, ptr(rowp), ptr(cols), root, iters,
                             ptr(out)))
    return out


def amd_order(rowp, cols):
    rowp, cols = _i32(rowp), _i32(cols)
    out = np.zeros(len(rowp) - 1, np.int32)
    check(lib().gf_amd_order(len(rowp) - 1, ptr(rowp), ptr(cols), ptr(out)))
    return out


def pattern_permute(rowp, cols, perm):
    rowp, cols, perm = _i32(rowp), _i32(cols), _i32(perm)
    n = len(rowp) - 1
    rp, cp = _PI(), _PI()
    check(lib().gf_pattern_permute(n, ptr(rowp), ptr(cols), ptr(perm),
                                   ctypes.byref(rp), ctypes.byref(cp)))
    return _take_host(rp, n + 1), _take_host(cp, int(rowp[-1]))


def iluk_symbolic(rowp, cols, lev_fill):
    rowp, cols = _i32(rowp), _i32(cols)
    n = len(rowp) - 1
    rp, cp, nnz = _PI(), _PI(), ctypes.c_int()
    check(lib().gf_iluk_symbolic(n, ptr(rowp), ptr(cols), lev_fill,
                                 ctypes.byref(rp), ctypes.byref(cp),
                                 ctypes.byref(nnz)))
    return _take_host(rp, n + 1), _take_host(cp, nnz.value)


def elem_ind_map(conn, rowp, cols, perm, nodes_per_elem=4):
    conn = _i32(conn).ravel()
    ne = len(conn) // nodes_per_elem
    out = np.zeros(ne * nodes_per_elem ** 2, np.int32)
    rowp, cols, perm = _i32(rowp), _i32(cols), _i32(perm)
    check(lib().gf_elem_ind_map(ne, nodes_per_elem, ptr(conn), ptr(rowp),
                                ptr(cols), ptr(perm), ptr(out)))
    return out


def transpose_map(rowp, cols):
    rowp, cols = _i32(rowp), _i32(cols)
    n = len(rowp) - 1
    a, b, c = _PI(), _PI(), _PI()
    check(lib().gf_transpose_map(n, ptr(rowp), ptr(cols), ctypes.byref(a),
                                 ctypes.byref(b), ctypes.byref(c)))
    nnz = int(rowp[-1])
    return _take_host(a, n + 1), _take_host(b, nnz), _take_host(c, nnz)


def bandwidth(rowp, cols):
    rowp, cols = _i32(rowp), _i32(cols)
    return lib().gf_bandwidth(len(rowp) - 1, ptr(rowp), ptr(cols))


def write_vtk(path, mesh, x=None):
    """Legacy-VTK file of a mesh and its displacements x[6*nnodes] in the
    unpermuted node order (printToVTK, include/mesh/vtk_writer.h:8-73)."""
    xp = np.ascontiguousarray(mesh.xpts, dtype=np.float64)
    cn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    xv = None
    if x is not None:
        xv = np.ascontiguousarray(x, dtype=np.float64)
        if xv.shape != (6 * mesh.nnodes,):
            raise ValueError(f"x must have {6 * mesh.nnodes} entries")
    check(lib().gf_write_vtk(os.fsencode(path), mesh.nnodes, ptr(xp),
                             mesh.nelems, ptr(cn),
                             None if xv is None else ptr(xv)))


class MeshLoader:
    """NASTRAN bulk-data reader with the reference's call surface
    (TACSMeshLoader, include/mesh/TACSMeshLoader.h:145-227): scan_bdf_file,
    then the node / element / BC arrays.  Host only."""

    def __init__(self):
        self.h = None

    def scan_bdf_file(self, path):
        """scanBDFFile (:763); raises GfError on an unreadable or bad file."""
        self.close()
        h = _P()
        code = lib().gf_bdf_read(os.fsencode(path), ctypes.byref(h))
        if code != 0:
            raise GfError(code, lib().gf_bdf_error().decode())
        self.h = h
        c = [ctypes.c_int() for _ in range(6)]
        check(lib().gf_bdf_sizes(self.h, *[ctypes.byref(v) for v in c]))
        (self.num_nodes, self.num_elements, self._conn_size, self._nbcs,
         self._nbc_vars, self.num_components) = [v.value for v in c]
        return self

    def connectivity(self):
        """getConnectivity (:184-187): xpts [nnodes, 3], elem_ptr, conn,
        elem_comp (property id - 1)."""
        x = np.zeros(3 * self.num_nodes)
        ep = np.zeros(self.num_elements + 1, np.int32)
        cn = np.zeros(self._conn_size, np.int32)
        cp = np.zeros(self.num_elements, np.int32)
        check(lib().gf_bdf_connectivity(self.h, ptr(x), ptr(ep), ptr(cn),
                                        ptr(cp)))
        return x.reshape(-1, 3), ep, cn, cp

    def bcs(self):
        """getBCs (:188-189): bc_nodes, bc_ptr, bc_vars, bc_vals."""
        bn = np.zeros(self._nbcs, np.int32)
        bp = np.zeros(self._nbcs + 1, np.int32)
        bv = np.zeros(self._nbc_vars, np.int32)
        bx = np.zeros(self._nbc_vars)
        check(lib().gf_bdf_bcs(self.h, ptr(bn), ptr(bp), ptr(bv), ptr(bx)))
        return bn, bp, bv, bx

    def assembler_data(self, vars_per_node=6):
        """getAssemblerCreatorData (:193-224): elem_conn, global BC DOFs
        (vars_per_node * node + var), xpts."""
        cn = np.zeros(self._conn_size, np.int32)
        b = np.zeros(self._nbc_vars, np.int32)
        x = np.zeros(3 * self.num_nodes)
        check(lib().gf_bdf_assembler_data(self.h, vars_per_node, ptr(cn),
                                          ptr(b), ptr(x)))
        return cn, b, x.reshape(-1, 3)

    def component_descript(self, comp):
        r = lib().gf_bdf_component_descript(self.h, comp)
        return None if r is None else r.decode()

    def element_descript(self, comp):
        r = lib().gf_bdf_element_descript(self.h, comp)
        return None if r is None else r.decode()

    def close(self):
        if self.h:
            lib().gf_bdf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# --------------------------------------------------------- device objects
class Context:
    """A device + HIP stream (gf_context)."""

    def __init__(self, device=0):
        import weakref

        h = _P()
        check(lib().gf_context_create(device, ctypes.byref(h)))
        self.h = h
        # objects whose device memory belongs to this context: closed first
        self._deps = weakref.WeakSet()

    def check(self, code):
        return check(code, self.h)

    def sync(self):
        self.check(lib().gf_context_synchronize(self.h))

    @property
    def stream(self):
        return lib().gf_context_stream(self.h)

    def close(self):
        if self.h:
            for d in list(self._deps):
                d.close()
            lib().gf_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ilu:
    """ILU analysis on a block pattern (gf_ilu)."""

    def __init__(self, ctx: Context, rowp, cols):
        self.ctx = ctx
        self.rowp, self.cols = _i32(rowp), _i32(cols)
        h = _P()
        ctx.check(lib().gf_ilu_create(ctx.h, len(rowp) - 1, ptr(self.rowp),
                                      ptr(self.cols), ctypes.byref(h)))
        self.h = h
        ctx._deps.add(self)

    def info(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().gf_ilu_info(self.h, ctypes.byref(a), ctypes.byref(b),
                          ctypes.byref(c))
        return {"nnzb": a.value, "levels_lower": b.value,
                "levels_upper": c.value}

    def factor(self, d_vals):
        zp = ctypes.c_int(-1)
        fn = (lib().gf_bsr_ilu_factor_f64 if d_vals.element_size() == 8
              else lib().gf_bsr_ilu_factor_f32)
        code = fn(self.h, ptr(d_vals), ctypes.byref(zp))
        if code != GF_OK:
            raise GfError(code, lib().gf_context_last_error(self.ctx.h).decode())
        return zp.value

    def apply(self, d_lu, d_b, d_tmp, d_x):
        fn = (lib().gf_bsr_ilu_apply_f64 if d_lu.element_size() == 8
              else lib().gf_bsr_ilu_apply_f32)
        self.ctx.check(fn(self.h, ptr(d_lu), ptr(d_b), ptr(d_tmp), ptr(d_x)))
        self.ctx.check(lib().gf_ilu_check(self.h))

    def close(self):
        if self.h:
            lib().gf_ilu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_options(**kw) -> Options:
    o = Options()
    lib().gf_options_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Problem:
    """A complete linear-static problem on the device (gf_problem)."""

    def __init__(self, ctx: Context, mesh, **opts):
        self.ctx = ctx
        self.mesh = mesh
        self.opt = default_options(**opts)
        conn = _i32(mesh.conn).ravel()
        xpts = np.ascontiguousarray(mesh.xpts, np.float64).ravel()
        data = np.ascontiguousarray(mesh.data, np.float64).ravel()
        bcs = _i32(mesh.bcs)
        h = _P()
        ctx.check(lib().gf_problem_create(
            ctx.h, mesh.nnodes, mesh.nelems, ptr(conn), ptr(xpts), ptr(data),
            len(bcs), ptr(bcs), ctypes.byref(self.opt), ctypes.byref(h)))
        self.h = h
        ctx._deps.add(self)

    def sizes(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(lib().gf_problem_sizes(self.h, ctypes.byref(a),
                                              ctypes.byref(b),
                                              ctypes.byref(c)))
        return a.value, b.value, c.value

    def device_arrays(self):
        ps = [_P() for _ in range(7)]
        self.ctx.check(lib().gf_problem_device_arrays(
            self.h, *[ctypes.byref(p) for p in ps]))
        return dict(zip(["krowp", "kcols", "kvals", "mrowp", "mcols", "mvals",
                         "perm"], [p.value for p in ps]))

    def ilu_handle(self):
        """The problem's gf_ilu (owned by the problem)."""
        return lib().gf_problem_ilu(self.h)

    def perm(self):
        out = np.zeros(self.mesh.nnodes, np.int32)
        self.ctx.check(lib().gf_problem_host_perm(self.h, ptr(out)))
        return out

    def assemble(self):
        self.ctx.check(lib().gf_problem_assemble(self.h))

    def factor(self):
        self.ctx.check(lib().gf_problem_factor(self.h))

    def solve(self, loads=None, x=None, raise_on_error=True):
        """Solve with host (numpy) or device (torch) loads; returns (x, info)."""
        loads = self.mesh.loads if loads is None else loads
        on_dev = not isinstance(loads, np.ndarray)
        if x is None:
            x = (np.zeros(6 * self.mesh.nnodes) if not on_dev
                 else loads.new_zeros(loads.shape))
        if not on_dev:
            loads = np.ascontiguousarray(loads, np.float64)
        info = SolveInfo()
        code = lib().gf_problem_solve(self.h, ptr(loads), ptr(x), int(on_dev),
                                      ctypes.byref(info))
        if code != GF_OK and raise_on_error:
            raise GfError(code, lib().gf_context_last_error(self.ctx.h).decode())
        return x, info

    def assemble_tangent(self, u, res=None):
        """Nonlinear tangent K(u) into the problem (with BCs) and, when res
        is given (host or device, like u), the residual R(u)."""
        on_dev = not isinstance(u, np.ndarray)
        if not on_dev:
            u = np.ascontiguousarray(u, np.float64)
        self.ctx.check(lib().gf_problem_assemble_tangent(
            self.h, ptr(u), ptr(res), int(on_dev)))
        return res

    def newton_solve(self, loads=None, u=None, max_newton=20,
                     newton_rtol=1e-8, raise_on_error=True):
        """Newton's method for the nonlinear shell, R(u) = loads; returns
        (u, info, history of ||f - R(u)|| / ||f||)."""
        loads = self.mesh.loads if loads is None else loads
        on_dev = not isinstance(loads, np.ndarray)
        if u is None:
            u = (np.zeros(6 * self.mesh.nnodes) if not on_dev
                 else loads.new_zeros(loads.shape))
        if not on_dev:
            loads = np.ascontiguousarray(loads, np.float64)
        hist = np.zeros(max_newton + 1)
        info = NewtonInfo()
        code = lib().gf_problem_newton_solve(
            self.h, ptr(loads), ptr(u), int(on_dev), max_newton, newton_rtol,
            ptr(hist), ctypes.byref(info))
        if code != GF_OK and raise_on_error:
            raise GfError(code, lib().gf_context_last_error(self.ctx.h).decode())
        return u, info, hist[:info.newton_iterations + 1]

    def close(self):
        if self.h:
            lib().gf_problem_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop (gf_server): FIFO admission under an HBM budget, worker
    slots with their own streams."""

    STATS = ["completed", "failed", "queued", "active", "p50_s", "p99_s",
             "peak_bytes", "elapsed_s", "escalated"]

    def __init__(self, budget_gb=200.0, slots=4, device=0):
        h = _P()
        check(lib().gf_server_create(device, budget_gb, slots, ctypes.byref(h)))
        self.h = h
        self.sizes = {}

    def submit(self, rid, mesh, loads=None, **opts):
        o = default_options(**opts)
        conn = _i32(mesh.conn).ravel()
        xpts = np.ascontiguousarray(mesh.xpts, np.float64).ravel()
        data = np.ascontiguousarray(mesh.data, np.float64).ravel()
        bcs = _i32(mesh.bcs)
        f = np.ascontiguousarray(mesh.loads if loads is None else loads,
                                 np.float64)
        self.sizes[rid] = mesh.nnodes
        return lib().gf_server_submit(self.h, rid, mesh.nnodes, mesh.nelems,
                                      ptr(conn), ptr(xpts), ptr(data),
                                      len(bcs), ptr(bcs), ptr(f),
                                      ctypes.byref(o))

    def poll(self):
        """(id, x, info, latency_s) of one finished request, or None."""
        rid = ctypes.c_longlong()
        info = SolveInfo()
        lat = ctypes.c_double()
        # size the output from the oldest finished id: poll into a buffer large
        # enough for any submitted request
        nmax = 6 * max(self.sizes.values()) if self.sizes else 0
        x = np.zeros(nmax)
        got = lib().gf_server_poll(self.h, ctypes.byref(rid), ptr(x), nmax,
                                   ctypes.byref(info), ctypes.byref(lat))
        if got != 1:
            return None
        n = 6 * self.sizes[rid.value]
        return rid.value, x[:n].copy(), info, lat.value

    def stats(self):
        out = np.zeros(9)
        check(lib().gf_server_stats(self.h, ptr(out)))
        return dict(zip(self.STATS, out.tolist()))

    def drain(self):
        check(lib().gf_server_drain(self.h))

    def close(self):
        if self.h:
            lib().gf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
