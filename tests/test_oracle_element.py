"""Pins for the oracle's MITC4 element (oracle/element.cpp).

Parity status: pinned.
  - the reference's 24x24 element stiffness fixture
    (examples/cusparse_solve/4_plate_soln.py:29-30) reproduced to its printed
    6 significant digits, with the connectivity order the reference used
  - symmetry, six rigid-body modes in the null space (flat and warped),
    R(u) = K u and U(u) = u.K.u/2, the reference's directional-derivative
    check of p.R against the energy (tests/shell_res/res_cs.cu:33-62)
  - membrane patch test on a distorted mesh, and convergence to the
    Kirchhoff closed form for a simply supported plate under a sine load
    (examples/plate_demo/plot_soln.py:150-160)
"""
import os

import numpy as np
import pytest

import oracle
from gpu_fem_amd import meshes as M

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DATA = (70e9, 0.3, 0.005, 0.0)

# warped element used by the reference's unit tests
# (tests/shell_res/local_utils.h:26: xpts[i] = 1.0345452 + 2.23123432 i + 0.323 i^2)
WARPED = np.array([1.0345452 + 2.23123432 * i + 0.323 * i * i for i in range(12)])


def sig6(x):
    out = np.zeros_like(x)
    nz = x != 0
    out[nz] = [float("%.5e" % v) for v in x[nz]]
    return out


def test_kelem_dense_fixture():
    ref = np.load(os.path.join(GOLDEN, "kelem_dense_plate1x0p5.npy"))
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 0.5, 0], [1, 0.5, 0]], float)
    # createPlateAssembler's counter-clockwise order {0, 1, 3, 2}
    order = [0, 1, 3, 2]
    K, _, _ = oracle.element(X[order], data=DATA)
    G = np.zeros((24, 24))
    for a in range(4):
        for b in range(4):
            G[6 * order[a]:6 * order[a] + 6, 6 * order[b]:6 * order[b] + 6] += \
                K[6 * a:6 * a + 6, 6 * b:6 * b + 6]
    # every entry matches the fixture at its printed precision
    np.testing.assert_array_equal(sig6(G), ref)
    # a plausible mistake -- tensor order instead of the generator's order --
    # does not match
    K2, _, _ = oracle.element(X, data=DATA)
    assert np.abs(sig6(K2) - ref).max() > 1e6


def rigid_modes(X):
    X = np.asarray(X).reshape(4, 3)
    modes = []
    for k in range(3):
        u = np.zeros(24)
        u[k::6] = 1.0
        modes.append(u)
    for k in range(3):
        th = np.zeros(3)
        th[k] = 1.0
        u = np.zeros(24)
        for a in range(4):
            u[6 * a:6 * a + 3] = np.cross(th, X[a])
            u[6 * a + 3:6 * a + 6] = th
        modes.append(u)
    return np.array(modes)


@pytest.mark.parametrize("geom", ["flat", "skew", "mild_warp", "ref_warped"])
def test_symmetry_and_rigid_modes(geom):
    if geom == "flat":
        X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
    elif geom == "skew":
        X = np.array([[0, 0, 0], [1.2, 0.1, 0], [0.3, 0.9, 0], [1.1, 1.3, 0]])
    elif geom == "mild_warp":
        X = np.array([[0, 0, 0], [1.1, 0.05, 0.02], [-0.1, 0.9, 0], [1, 1, -0.08]])
    else:
        X = WARPED.reshape(4, 3)
    K, _, _ = oracle.element(X, data=DATA)
    scale = np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-12 * scale
    for r in rigid_modes(X):
        assert np.abs(K @ r).max() <= 1e-9 * scale * np.abs(r).max()
    # no spurious mechanism on reasonable shapes: exactly 6 zero-energy modes
    # (the reference's test element is extremely distorted and has a 7th
    # soft mode, so only the rigid modes are checked there)
    if geom != "ref_warped":
        w = np.sort(np.abs(np.linalg.eigvalsh(0.5 * (K + K.T))))
        assert w[5] < 1e-9 * w[6]


@pytest.mark.parametrize("geom", ["flat", "warped"])
def test_residual_and_energy_consistent(geom):
    X = (np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)
         if geom == "flat" else WARPED.reshape(4, 3))
    rng = np.random.default_rng(1)
    u = rng.standard_normal(24) * 1e-3
    K, R, U = oracle.element(X, u, DATA)
    np.testing.assert_allclose(R, K @ u, rtol=0, atol=1e-12 * np.abs(K).max() * 1e-3)
    assert abs(U - 0.5 * u @ K @ u) <= 1e-12 * abs(U)
    # directional derivative of the energy along p equals p.R
    p = rng.standard_normal(24)
    h = 1e-6
    _, _, Up = oracle.element(X, u + h * p, DATA)
    _, _, Um = oracle.element(X, u - h * p, DATA)
    fd = (Up - Um) / (2 * h)
    assert abs(fd - p @ R) <= 1e-7 * abs(p @ R)


def test_reference_unit_inputs():
    # the reference's one-element inputs (tests/shell_jac/jac.cu:36-38)
    X = WARPED.reshape(4, 3)
    v = np.array([(1.4543 + 6.4323 * i) * 1e-6 for i in range(24)])
    p = np.array([-1.4543 + 2.312 * 6.4323 * i for i in range(24)])
    K, R, U = oracle.element(X, v, DATA)
    # p.R is the energy's directional derivative; R is linear in v
    K2, R2, _ = oracle.element(X, 2 * v, DATA)
    np.testing.assert_allclose(R2, 2 * R, rtol=1e-12, atol=1e-12 * np.abs(R).max())
    assert np.isfinite(p @ R) and U > 0


def _dense_system(mesh):
    n = mesh.nnodes
    rp, cp = oracle.node_graph(mesh.conn, n)
    perm = np.arange(n, dtype=np.int32)
    emap = oracle.elem_ind_map(mesh.conn, rp, cp, perm)
    kv, _ = oracle.assemble_jacobian(mesh.conn, mesh.xpts, mesh.data, emap,
                                     perm, int(rp[-1]), with_res=False)
    K = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            j = cp[k]
            K[6 * i:6 * i + 6, 6 * j:6 * j + 6] = kv[36 * k:36 * k + 36].reshape(6, 6)
    return K


def test_membrane_patch_parallelograms():
    # constant membrane strain on a mesh of parallelograms rotated into 3-D:
    # interior nodes carry no force.  (The tied MITC4 membrane strains are
    # exact for parallelograms; on general quadrilaterals the covariant
    # strains vary quadratically and the patch test is not passed exactly.)
    mesh = M.plate(4, 4, 1.0, 1.0)
    xp = mesh.xpts.copy()
    inner = [i for i in range(mesh.nnodes)
             if 0 < xp[i, 0] < 1 and 0 < xp[i, 1] < 1]
    xp = xp @ np.array([[1.2, 0.3, 0], [-0.2, 0.9, 0], [0, 0, 1]]).T
    R = np.linalg.qr(np.random.default_rng(0).standard_normal((3, 3)))[0]
    mesh.xpts = xp @ R.T
    K = _dense_system(mesh)
    u = np.zeros((mesh.nnodes, 6))
    u2 = np.zeros((mesh.nnodes, 3))
    u2[:, 0] = 1e-3 * xp[:, 0] + 3e-4 * xp[:, 1]
    u2[:, 1] = -4e-4 * xp[:, 0] + 2e-3 * xp[:, 1]
    u[:, :3] = u2 @ R.T
    # drilling rotation equal to the in-plane rotation (dv/dx - du/dy)/2
    u[:, 3:] = np.array([0, 0, 0.5 * (-4e-4 - 3e-4)]) @ R.T
    f = (K @ u.ravel()).reshape(-1, 6)
    assert np.abs(f[inner]).max() <= 1e-12 * np.abs(f).max()
    # a wrong drilling rotation is detected (the drill strain is active)
    u[:, 3:] = 0.0
    f = (K @ u.ravel()).reshape(-1, 6)
    assert np.abs(f[inner]).max() > 1e-3 * np.abs(f).max()


def test_kirchhoff_closed_form():
    # simply supported square plate, sine load, thin: w_max -> Q/(D pi^4 4/a^4)
    a, t, E, nu, Q = 1.0, 0.005, 70e9, 0.3, 1e3
    mesh = M.plate(16, 16, a, a, E=E, nu=nu, thick=t, bc="ss", load="sine",
                   load_mag=Q)
    K = _dense_system(mesh)
    for b in mesh.bcs:
        K[b, :] = 0
        K[:, b] = 0
        K[b, b] = 1
    x = oracle.dense_solve(K, mesh.loads)
    D = E * t ** 3 / (12 * (1 - nu ** 2))
    w_exact = Q / (D * np.pi ** 4 * (2 / a ** 2) ** 2)
    w_max = np.abs(x[2::6]).max()
    assert abs(w_max - w_exact) / w_exact < 0.02
