"""Seeded mesh, boundary-condition and load generators for the linear-static
shell workloads (host side, numpy).

Reference: ``examples/plate_demo/_plate_utils.h:9-144`` (createPlateAssembler),
``:146-184`` (getPlateLoads), ``:186-212`` (getPlatePointLoad).  The reference
has no cylinder or wingbox generator; those are defined here.

Connectivity is emitted in the basis' tensor order {(-1,-1), (1,-1), (-1,1),
(1,1)} (``include/shell/basis.h:138-142``).  ``ccw=True`` reproduces the
reference plate generator's counter-clockwise order {n, n+1, n+nnx+1, n+nnx}
(``_plate_utils.h:103-104``), which the reference's own fixtures were made
with.

Every generator returns a :class:`Mesh`; element data rows are
``(E, nu, thick, tOffset)``; BCs are DOF indices ``6*node + k`` in the
unpermuted node numbering; loads have ``6*nnodes`` entries.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

RT3 = 0.577350269189626  # include/shell/quadrature.h:9
REF_PI = 3.1415926535897  # examples/plate_demo/_plate_utils.h:165


@dataclass
class Mesh:
    conn: np.ndarray       # int32 [nelems, 4]
    xpts: np.ndarray       # float64 [nnodes, 3]
    data: np.ndarray       # float64 [nelems, 4]
    bcs: np.ndarray        # int32 [nbcs]
    loads: np.ndarray      # float64 [6*nnodes]
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nnodes(self) -> int:
        return self.xpts.shape[0]

    @property
    def nelems(self) -> int:
        return self.conn.shape[0]


def _data(nelems, E, nu, thick, t_offset=0.0):
    d = np.empty((nelems, 4))
    d[:, 0] = E
    d[:, 1] = nu
    d[:, 2] = thick
    d[:, 3] = t_offset
    return d


def plate_conn(nxe, nye, ccw=False):
    nnx = nxe + 1
    iy, ix = np.meshgrid(np.arange(nye), np.arange(nxe), indexing="ij")
    n0 = (nnx * iy + ix).ravel()
    if ccw:
        conn = np.stack([n0, n0 + 1, n0 + nnx + 1, n0 + nnx], axis=1)
    else:
        conn = np.stack([n0, n0 + 1, n0 + nnx, n0 + nnx + 1], axis=1)
    return conn.astype(np.int32)


def plate_xpts(nxe, nye, Lx, Ly):
    nnx, nny = nxe + 1, nye + 1
    dx, dy = Lx / nxe, Ly / nye
    iy, ix = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    x = np.zeros((nnx * nny, 3))
    x[:, 0] = (dx * ix).ravel()
    x[:, 1] = (dy * iy).ravel()
    return x


def plate_bcs_ss_reference(nxe, nye):
    """Simply supported pattern of ``_plate_utils.h:51-82`` (same order)."""
    nnx, nny = nxe + 1, nye + 1
    bcs = list(range(6))
    for ix in range(1, nnx):
        n = ix
        bcs += [6 * n + 1, 6 * n + 2]
    for iy in range(1, nny):
        n = nnx * iy
        bcs += [6 * n, 6 * n + 2]
        n = nnx * iy + nnx - 1
        bcs += [6 * n + 2]
    for ix in range(1, nnx - 1):
        n = nnx * (nny - 1) + ix
        bcs += [6 * n + 2]
    return np.array(bcs, np.int32)


def plate_bcs_clamped(nxe, nye):
    nnx, nny = nxe + 1, nye + 1
    iy, ix = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    edge = ((ix == 0) | (ix == nnx - 1) | (iy == 0) | (iy == nny - 1)).ravel()
    nodes = np.nonzero(edge)[0]
    return (6 * nodes[:, None] + np.arange(6)[None, :]).ravel().astype(np.int32)


def plate_loads_nodal_reference(nxe, nye, Lx, Ly, load_mag):
    """Point-sampled sine load of ``_plate_utils.h:146-184`` (w DOF only)."""
    nnx, nny = nxe + 1, nye + 1
    dx, dy = Lx / nxe, Ly / nye
    iy, ix = np.meshgrid(np.arange(nny), np.arange(nnx), indexing="ij")
    x, y = ix * dx, iy * dy
    f = np.zeros(6 * nnx * nny)
    f[2::6] = (load_mag * np.sin(REF_PI * x / Lx) *
               np.sin(REF_PI * y / Ly)).ravel()
    return f


def consistent_pressure_loads(conn, xpts, pressure, normal_fn=None, order=3):
    """Nodal forces f_a = int p(x) N_a n dA over every element.

    ``pressure(x)`` gives the pressure at points x [m, 3]; the force acts
    along the element normal (``normal_fn`` may override it).  Gauss order
    ``order`` per direction on the bilinear map (tensor node order).
    """
    g, w = np.polynomial.legendre.leggauss(order)
    X = xpts[conn]  # [ne, 4, 3]
    f = np.zeros((xpts.shape[0], 3))
    xa = np.array([-1.0, 1.0, -1.0, 1.0])
    ea = np.array([-1.0, -1.0, 1.0, 1.0])
    for i, xi in enumerate(g):
        for j, eta in enumerate(g):
            N = 0.25 * (1 + xa * xi) * (1 + ea * eta)
            Nxi = 0.25 * xa * (1 + ea * eta)
            Neta = 0.25 * ea * (1 + xa * xi)
            x = np.einsum("a,eak->ek", N, X)
            Xxi = np.einsum("a,eak->ek", Nxi, X)
            Xeta = np.einsum("a,eak->ek", Neta, X)
            nrm = np.cross(Xxi, Xeta)  # |nrm| = dA / (dxi deta)
            if normal_fn is not None:
                dA = np.linalg.norm(nrm, axis=1)
                nrm = normal_fn(x) * dA[:, None]
            p = pressure(x)
            for a in range(4):
                np.add.at(f, conn[:, a], (w[i] * w[j] * N[a] * p)[:, None] * nrm)
    out = np.zeros(6 * xpts.shape[0])
    out[0::6], out[1::6], out[2::6] = f[:, 0], f[:, 1], f[:, 2]
    return out


def plate(nxe, nye, Lx=1.0, Ly=1.0, E=70e9, nu=0.3, thick=0.01,
          bc="clamped", load="uniform", load_mag=1e3, ccw=False):
    """Flat plate in the z = 0 plane.

    bc: "clamped" (all 6 DOFs on every edge node) or "ss" (the reference's
    simply supported pattern).  load: "uniform" (consistent, magnitude
    load_mag Pa), "sine" (consistent load_mag sin(pi x/Lx) sin(pi y/Ly)),
    "uniform_sine" (consistent load_mag (1 + sin(pi x/Lx) sin(pi y/Ly)) / 2)
    or "nodal_sine" (the reference's point-sampled nodal values).
    """
    conn = plate_conn(nxe, nye, ccw=ccw)
    xpts = plate_xpts(nxe, nye, Lx, Ly)
    bcs = (plate_bcs_clamped(nxe, nye) if bc == "clamped"
           else plate_bcs_ss_reference(nxe, nye))
    if load == "uniform":
        loads = consistent_pressure_loads(conn, xpts,
                                          lambda x: np.full(len(x), load_mag))
    elif load == "sine":
        loads = consistent_pressure_loads(
            conn, xpts, lambda x: load_mag * np.sin(np.pi * x[:, 0] / Lx) *
            np.sin(np.pi * x[:, 1] / Ly))
    elif load == "uniform_sine":
        loads = consistent_pressure_loads(
            conn, xpts, lambda x: 0.5 * load_mag * (
                1.0 + np.sin(np.pi * x[:, 0] / Lx) * np.sin(np.pi * x[:, 1] / Ly)))
    elif load == "nodal_sine":
        loads = plate_loads_nodal_reference(nxe, nye, Lx, Ly, load_mag)
    else:
        raise ValueError(load)
    loads[bcs] = 0.0
    return Mesh(conn, xpts, _data(len(conn), E, nu, thick), bcs, loads,
                name=f"plate{nxe}x{nye}",
                meta=dict(nxe=nxe, nye=nye, Lx=Lx, Ly=Ly, bc=bc, load=load))


def cylinder(nc, na, R=1.0, L=3.0, E=70e9, nu=0.3, t_range=(1e-3, 8e-3),
             seed=0, pressure=50e3, axial_line_load=10e3, bc="clamped"):
    """Open cylinder, nc circumferential x na axial quads (periodic in theta).

    Nodes: nc around x (na + 1) along z.  The z = 0 end is clamped (bc
    "clamped", all 6 DOFs) or pinned (bc "ss", translations only); internal
    pressure acts outward (consistent loads) and a uniform axial line load
    (N/m) acts on the z = L end ring.  Per-element thickness is uniform in
    t_range from ``seed`` (numpy PCG64).
    """
    th = 2 * np.pi * np.arange(nc) / nc
    z = L * np.arange(na + 1) / na
    ZZ, TT = np.meshgrid(z, th, indexing="ij")
    xpts = np.stack([R * np.cos(TT).ravel(), R * np.sin(TT).ravel(),
                     ZZ.ravel()], axis=1)
    ia, ic = np.meshgrid(np.arange(na), np.arange(nc), indexing="ij")
    n00 = (nc * ia + ic).ravel()
    n10 = (nc * ia + (ic + 1) % nc).ravel()
    conn = np.stack([n00, n10, n00 + nc, n10 + nc], axis=1).astype(np.int32)
    rng = np.random.default_rng(seed)
    thick = rng.uniform(t_range[0], t_range[1], size=len(conn))
    data = _data(len(conn), E, nu, 1.0)
    data[:, 2] = thick
    ndof = 6 if bc == "clamped" else 3
    bcs = (6 * np.arange(nc)[:, None] + np.arange(ndof)[None, :]).ravel()
    radial = lambda x: np.stack([x[:, 0], x[:, 1], 0 * x[:, 0]], 1) / R
    loads = consistent_pressure_loads(conn, xpts,
                                      lambda x: np.full(len(x), pressure),
                                      normal_fn=radial)
    end = nc * na + np.arange(nc)
    loads[6 * end + 2] += axial_line_load * (2 * np.pi * R / nc)
    loads[bcs] = 0.0
    return Mesh(conn, xpts, data, bcs.astype(np.int32), loads,
                name=f"cylinder{nc}x{na}",
                meta=dict(nc=nc, na=na, R=R, L=L, seed=seed, bc=bc))


def wingbox(nc=120, nh=16, nsb=30, nribs=50, span=30.0, c_root=6.0,
            c_tip=1.5, h_frac=0.12, E=70e9, nu=0.3, t_range=(2e-3, 20e-3),
            seed=0, p_root=20e3):
    """Synthetic tapered wingbox from structured quad patches that share
    their edge nodes.

    The span runs along y in [0, span]; the chord c(y) tapers linearly from
    c_root to c_tip between the front spar (x = 0) and the rear spar
    (x = c(y)); the box height is h(y) = h_frac c(y) with the lower and upper
    skins at z = -h/2 and +h/2.  ``nribs`` ribs at equal spacing (the first
    at the root, the last at the tip) divide the span into nribs - 1 bays of
    ``nsb`` elements; skins have ``nc`` elements along the chord, spars and
    ribs ``nh`` through the height.  Every panel (one bay of a skin or a
    spar, or one rib) has its own thickness, uniform in ``t_range`` from
    ``seed`` (numpy PCG64).  The root (y = 0) is clamped (all 6 DOFs); an
    upward elliptic pressure p(y) = p_root sqrt(1 - (y/span)^2) acts on the
    lower skin as consistent nodal loads.  The defaults give 489,362 nodes
    and 495,840 elements.
    """
    Ns = (nribs - 1) * nsb
    y = span * np.arange(Ns + 1) / Ns
    c = c_root + (c_tip - c_root) * y / span
    h = h_frac * c
    nskin = (nc + 1) * (Ns + 1)
    nspar = (nh - 1) * (Ns + 1)
    nrib = (nc - 1) * (nh - 1)
    offU, offF = nskin, 2 * nskin
    offR = offF + nspar
    offB = offR + nspar
    nnodes = offB + nribs * nrib

    def skin(i, j, upper):
        return (offU if upper else 0) + i + (nc + 1) * j

    def spar(k, j, rear):
        k, j = np.broadcast_arrays(k, j)
        i = nc if rear else 0
        out = (offR if rear else offF) + (k - 1) + (nh - 1) * j
        out = np.where(k == 0, skin(i, j, False), out)
        return np.where(k == nh, skin(i, j, True), out)

    def rib(r, i, k):
        i, k = np.broadcast_arrays(i, k)
        jr = r * nsb
        out = offB + r * nrib + (i - 1) + (nc - 1) * (k - 1)
        out = np.where(i == 0, spar(k, jr, False), out)
        out = np.where(i == nc, spar(k, jr, True), out)
        out = np.where(k == 0, skin(i, jr, False), out)
        return np.where(k == nh, skin(i, jr, True), out)

    xpts = np.zeros((nnodes, 3))
    I, J = np.meshgrid(np.arange(nc + 1), np.arange(Ns + 1), indexing="ij")
    for upper in (False, True):
        idx = skin(I, J, upper)
        xpts[idx, 0] = c[J] * I / nc
        xpts[idx, 1] = y[J]
        xpts[idx, 2] = (0.5 if upper else -0.5) * h[J]
    K, J = np.meshgrid(np.arange(1, nh), np.arange(Ns + 1), indexing="ij")
    for rear in (False, True):
        idx = spar(K, J, rear)
        xpts[idx, 0] = c[J] if rear else 0.0
        xpts[idx, 1] = y[J]
        xpts[idx, 2] = h[J] * (K / nh - 0.5)
    I, K = np.meshgrid(np.arange(1, nc), np.arange(1, nh), indexing="ij")
    for r in range(nribs):
        jr = r * nsb
        idx = rib(r, I, K)
        xpts[idx, 0] = c[jr] * I / nc
        xpts[idx, 1] = y[jr]
        xpts[idx, 2] = h[jr] * (K / nh - 0.5)

    conns, panels = [], []
    # skins: lower then upper, panel = bay
    I, J = np.meshgrid(np.arange(nc), np.arange(Ns), indexing="ij")
    for s_id, upper in enumerate((False, True)):
        q = [skin(I, J, upper), skin(I + 1, J, upper), skin(I, J + 1, upper),
             skin(I + 1, J + 1, upper)]
        conns.append(np.stack([a.ravel() for a in q], 1))
        panels.append((s_id * (nribs - 1) + J // nsb).ravel())
    nlower = nc * Ns
    # spars: front then rear, panel = bay
    K, J = np.meshgrid(np.arange(nh), np.arange(Ns), indexing="ij")
    for s_id, rear in enumerate((False, True)):
        q = [spar(K, J, rear), spar(K, J + 1, rear), spar(K + 1, J, rear),
             spar(K + 1, J + 1, rear)]
        conns.append(np.stack([a.ravel() for a in q], 1))
        panels.append(((2 + s_id) * (nribs - 1) + J // nsb).ravel())
    # ribs: panel = rib
    I, K = np.meshgrid(np.arange(nc), np.arange(nh), indexing="ij")
    for r in range(nribs):
        q = [rib(r, I, K), rib(r, I + 1, K), rib(r, I, K + 1),
             rib(r, I + 1, K + 1)]
        conns.append(np.stack([a.ravel() for a in q], 1))
        panels.append(np.full(nc * nh, 4 * (nribs - 1) + r))
    conn = np.concatenate(conns).astype(np.int32)
    panel = np.concatenate(panels)
    rng = np.random.default_rng(seed)
    t_panel = rng.uniform(t_range[0], t_range[1], size=5 * nribs - 4)
    data = _data(len(conn), E, nu, 1.0)
    data[:, 2] = t_panel[panel]
    root = np.flatnonzero(np.abs(xpts[:, 1]) < 1e-12 * span)
    bcs = (6 * root[:, None] + np.arange(6)[None, :]).ravel()
    up = lambda x: np.tile(np.array([0.0, 0.0, 1.0]), (len(x), 1))
    loads = consistent_pressure_loads(
        conn[:nlower], xpts,
        lambda x: p_root * np.sqrt(np.maximum(0.0, 1.0 - (x[:, 1] / span) ** 2)),
        normal_fn=up)
    loads[bcs] = 0.0
    return Mesh(conn, xpts, data, bcs.astype(np.int32), loads,
                name=f"wingbox{nc}x{nh}x{Ns}r{nribs}",
                meta=dict(nc=nc, nh=nh, nsb=nsb, nribs=nribs, span=span,
                          c_root=c_root, c_tip=c_tip, h_frac=h_frac, seed=seed,
                          panels=5 * nribs - 4))


def from_bdf(path, E=70e9, nu=0.3, thick=0.005, loads=None, name=None):
    """Shell mesh from a NASTRAN bulk-data file, one material and thickness
    for every element (ElementAssembler::createFromBDF with one Data,
    include/assembler.h:92-115; the uCRM drivers use E = 70 GPa, nu = 0.3,
    t = 5 mm, examples/uCRM/linear_static.cu:31-34).  Every element must be
    a 4-node quad (CQUAD4 / CQUADR); BCs are the SPC DOFs as 6*node + var.
    ``loads``: None (zero), an array of 6*nnodes, or a callable
    f(conn, xpts) returning one."""
    from . import MeshLoader

    ml = MeshLoader().scan_bdf_file(path)
    try:
        xpts, eptr, conn, _ = ml.connectivity()
        if ml.num_elements and not np.all(np.diff(eptr) == 4):
            bad = int(np.argmax(np.diff(eptr) != 4))
            raise ValueError(f"{path}: element {bad} (in id order) is not a "
                             f"4-node quad ({ml.element_descript(0)} ...)")
        _, _, bv, _ = ml.bcs()
        if np.any(bv >= 6):
            raise ValueError(f"{path}: SPC component digits 7-8 do not exist "
                             "for a 6-DOF shell node")
        _, bcs, _ = ml.assembler_data(6)
    finally:
        ml.close()
    conn = conn.reshape(-1, 4).astype(np.int32)
    nn = xpts.shape[0]
    if loads is None:
        f = np.zeros(6 * nn)
    elif callable(loads):
        f = np.asarray(loads(conn, xpts), dtype=np.float64)
    else:
        f = np.asarray(loads, dtype=np.float64).copy()
    if f.shape != (6 * nn,):
        raise ValueError(f"loads must have {6 * nn} entries")
    return Mesh(conn=conn, xpts=np.ascontiguousarray(xpts),
                data=_data(conn.shape[0], E, nu, thick), bcs=bcs.astype(np.int32),
                loads=f, name=name or os.path.splitext(os.path.basename(path))[0])


def to_bdf(mesh, path, large_field=False):
    """Write a mesh's nodes, CQUAD4 elements (property 1) and SPCs as NASTRAN
    bulk data, small (8-column) or large (16-column) field: from_bdf reads
    back the same connectivity and BCs, and the coordinates rounded to 10
    significant digits (large field) or 5 (small field)."""
    lines = ["BEGIN BULK"]
    for i, (x, y, z) in enumerate(mesh.xpts):
        if large_field:
            lines.append(f"GRID*   {i + 1:<16d}{0:<16d}{x:<16.9E}{y:<16.9E}*")
            lines.append(f"*       {z:<16.9E}")
        else:
            lines.append(f"GRID    {i + 1:<8d}{0:<8d}{_f8(x)}{_f8(y)}{_f8(z)}")
    for e, c in enumerate(mesh.conn):
        g = [c[0] + 1, c[1] + 1, c[3] + 1, c[2] + 1]  # around the element
        lines.append(f"CQUAD4  {e + 1:<8d}{1:<8d}" + "".join(f"{v:<8d}" for v in g))
    by_node = {}
    for d in np.asarray(mesh.bcs):
        by_node.setdefault(int(d) // 6, []).append(int(d) % 6 + 1)
    for nd, comps in sorted(by_node.items()):
        cs = "".join(str(v) for v in sorted(set(comps)))
        lines.append(f"SPC     {1:<8d}{nd + 1:<8d}{cs:<8s}{0.0:<8.1f}")
    lines.append("ENDDATA")
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _f8(v):
    """A real in 8 columns (NASTRAN small field)."""
    for p in range(7, -1, -1):
        s = f"{v:.{p}E}".replace("E+0", "+").replace("E-0", "-").replace("E+", "+").replace("E-", "-")
        if len(s) <= 8:
            return f"{s:<8s}"
    raise ValueError(f"{v} does not fit 8 columns")


def served_requests(n, seed=0, scale=1.0):
    """The served request mix: seeded plates (n x n quads, n log-uniform in
    [32, 256]) and cylinders (nc x na from 64 x 128 to 192 x 384,
    log-uniform), with probability 1/2 each; thickness uniform in [1, 10] mm;
    uniform-plus-sinusoidal pressure with magnitude uniform in [0.5, 5] kPa;
    clamped or simply supported with probability 1/2.  ``scale`` < 1 shrinks
    every mesh side (for quick runs)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = float(rng.uniform(1e-3, 10e-3))
        p0 = float(rng.uniform(500.0, 5000.0))
        bc = "clamped" if rng.random() < 0.5 else "ss"
        if rng.random() < 0.5:
            m = int(round(np.exp(rng.uniform(np.log(32), np.log(256))) * scale))
            m = max(m, 4)
            mesh = plate(m, m, 1.0, 1.0, thick=t, bc=bc, load="uniform_sine",
                         load_mag=p0)
        else:
            f = np.exp(rng.uniform(np.log(1.0), np.log(3.0)))
            nc = max(int(round(64 * f * scale)), 8)
            na = max(int(round(128 * f * scale)), 4)
            mesh = cylinder(nc, na, t_range=(t, t), seed=int(rng.integers(1 << 30)),
                            pressure=p0, axial_line_load=0.0, bc=bc)
        mesh.meta["request"] = i
        out.append(mesh)
    return out
