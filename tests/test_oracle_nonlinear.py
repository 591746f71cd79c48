"""Oracle of the nonlinear shell (IsotropicShell<..., isNonlinear = true>,
include/shell/physics.h:6-17; strain include/shell/a2d/a2dshellstrain.h:72-93).

Pins, independent of the formulas inside the oracle:
- the local displacement gradients u0x, u1x and the nonlinear bending strains
  for linear displacement and rotation fields on a rectangle, where the
  gradients are known by hand (T = I, Xd^-1 T = diag(2/a, 2/b, 1), Z = 0,
  d = q x e_z), also with the element and the fields rotated into 3-D;
- R = dU/du and K = dR/du by central differences of the oracle's own energy
  and residual (flat, skew and warped elements);
- the linear limit (K(0) is the linear element's K) and symmetry of K.
"""
import numpy as np
import pytest

import oracle

DATA = (70e9, 0.3, 0.005, 0.0)


def rect(a=1.0, b=0.5):
    # tensor order (basis.h:138-142): (0,0) (a,0) (0,b) (a,b)
    return np.array([0, 0, 0, a, 0, 0, 0, b, 0, a, b, 0], float)


def rotation(seed=3):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def linear_fields(X, Gu, U0, Gq, Q0, R=np.eye(3)):
    """Nodal DOFs of u(x) = U0 + Gu [x, y], q(x) = Q0 + Gq [x, y] given in the
    element's own (x, y, z) frame, rotated by R into global components."""
    v = np.zeros(24)
    for a in range(4):
        xy = X[3 * a:3 * a + 2]
        v[6 * a:6 * a + 3] = R @ (U0 + Gu @ xy)
        v[6 * a + 3:6 * a + 6] = R @ (Q0 + Gq @ xy)
    return v


@pytest.mark.parametrize("rotated", [False, True])
def test_nonlinear_strain_known_gradients(rotated):
    a, b = 1.0, 0.5
    X0 = rect(a, b)
    rng = np.random.default_rng(7)
    Gu = 1e-2 * rng.standard_normal((3, 2))
    U0 = 1e-2 * rng.standard_normal(3)
    Gq = 1e-2 * rng.standard_normal((3, 2))
    Q0 = 1e-2 * rng.standard_normal(3)
    R = rotation() if rotated else np.eye(3)
    X = np.concatenate([R @ X0[3 * i:3 * i + 3] for i in range(4)])
    v = linear_fields(X0, Gu, U0, Gq, Q0, R)
    rt3 = 0.577350269189626
    for iq in range(4):
        xi = rt3 if iq % 2 else -rt3
        eta = rt3 if iq // 2 else -rt3
        x, y = 0.5 * a * (1 + xi), 0.5 * b * (1 + eta)
        q = Q0 + Gq @ [x, y]
        # d = q x e_z = (q_y, -q_x, 0); its gradient from Gq
        d = np.array([q[1], -q[0], 0.0])
        dd = np.array([Gq[1], -Gq[0], [0.0, 0.0]])  # d(d_r)/d(x, y)
        u0x = np.column_stack([Gu[:, 0], Gu[:, 1], d])
        u1x = np.column_stack([dd[:, 0], dd[:, 1], np.zeros(3)])
        E, o0, o1, _ = oracle.strain_nl(iq, X, v)
        # the local frame of the rotated element is R's image of e_x, e_y,
        # e_z, so the local gradients are those of the flat element
        assert np.allclose(o0[:, :2], u0x[:, :2], rtol=0, atol=1e-14)
        assert np.allclose(o0[:, 2], u0x[:, 2], rtol=0, atol=1e-14)
        assert np.allclose(o1[:, :2], u1x[:, :2], rtol=0, atol=1e-14)
        # bending strains of NonlinearShellStrainCore (a2dshellstrain.h:80-85)
        k11 = u1x[0, 0] + u0x[:, 0] @ u1x[:, 0]
        k22 = u1x[1, 1] + u0x[:, 1] @ u1x[:, 1]
        k12 = (u1x[0, 1] + u1x[1, 0] + u0x[:, 0] @ u1x[:, 1]
               + u1x[:, 0] @ u0x[:, 1])
        assert np.allclose(E[3:6], [k11, k22, k12], rtol=1e-12, atol=1e-16)
        # the other strains are the linear element's
        El, _ = oracle.strain(iq, X, v)
        assert np.array_equal(E[[0, 1, 2, 6, 7, 8]], El[[0, 1, 2, 6, 7, 8]])


def elements():
    skew = np.array([0, 0, 0, 1.1, 0.1, 0, 0.2, 0.6, 0, 1.25, 0.75, 0], float)
    warped = skew.copy()
    warped[11] = 0.08
    curved = np.array([1.0345452 + 2.23123432 * i + 0.323 * i * i
                       for i in range(12)])  # tests/shell_res/local_utils.h:26
    return {"rect": rect(), "skew": skew, "warped": warped, "curved": curved}


@pytest.mark.parametrize("name", list(elements()))
def test_nonlinear_residual_and_tangent_are_derivatives(name):
    X = elements()[name]
    rng = np.random.default_rng(11)
    # large enough that the quadratic terms matter
    u = 2e-2 * rng.standard_normal(24)
    K, R, U = oracle.element_nl(X, u, DATA)
    h = 1e-6
    g = np.zeros(24)
    Kfd = np.zeros((24, 24))
    for j in range(24):
        e = np.zeros(24)
        e[j] = h
        _, Rp, Up = oracle.element_nl(X, u + e, DATA)
        _, Rm, Um = oracle.element_nl(X, u - e, DATA)
        g[j] = (Up - Um) / (2 * h)
        Kfd[:, j] = (Rp - Rm) / (2 * h)
    assert np.abs(g - R).max() <= 1e-7 * np.abs(R).max()
    assert np.abs(Kfd - K).max() <= 1e-7 * np.abs(K).max()
    assert np.abs(K - K.T).max() <= 1e-14 * np.abs(K).max()
    # the nonlinear terms are exercised: K(u) differs from K(0) by a fair
    # fraction of the bending rigidity E t^3 / (12 (1 - nu^2)) (the membrane
    # and drill entries of K are 1e5 times larger and unchanged)
    K0, _, _ = oracle.element(X, None, DATA)
    E, nu, t = DATA[0], DATA[1], DATA[2]
    assert np.abs(K - K0).max() > 1e-2 * E * t ** 3 / (12 * (1 - nu ** 2))


@pytest.mark.parametrize("name", list(elements()))
def test_nonlinear_linear_limit(name):
    X = elements()[name]
    K0, _, _ = oracle.element(X, None, DATA)
    Kn, Rn, Un = oracle.element_nl(X, np.zeros(24), DATA)
    assert np.array_equal(Kn, K0)
    assert not Rn.any() and Un == 0.0
    rng = np.random.default_rng(2)
    u = rng.standard_normal(24)
    eps = 1e-9
    _, Re, _ = oracle.element_nl(X, eps * u, DATA)
    assert np.allclose(Re / eps, K0 @ u, rtol=1e-6,
                       atol=1e-6 * np.abs(K0 @ u).max())


def test_nonlinear_assembly_matches_element_sum():
    from gpu_fem_amd import meshes as M

    m = M.plate(4, 3, 1.0, 0.75, thick=0.01, bc="clamped", load="uniform")
    rp, cp = oracle.node_graph(m.conn, m.nnodes)
    perm = np.arange(m.nnodes, dtype=np.int32)
    emap = oracle.elem_ind_map(m.conn, rp, cp, perm)
    rng = np.random.default_rng(5)
    u = 1e-2 * rng.standard_normal(6 * m.nnodes)
    kv, res, U = oracle.assemble_tangent(m.conn, m.xpts, m.data, emap, perm,
                                         int(rp[-1]), u)
    # dense sum of element tangents
    n = 6 * m.nnodes
    Kd, Rd, Ud = np.zeros((n, n)), np.zeros(n), 0.0
    conn = np.asarray(m.conn).reshape(-1, 4)
    xp = np.asarray(m.xpts).ravel()
    dat = np.asarray(m.data).reshape(-1, 4)
    for e in range(m.nelems):
        nd = conn[e]
        idx = np.concatenate([6 * k + np.arange(6) for k in nd])
        Xe = np.concatenate([xp[3 * k:3 * k + 3] for k in nd])
        Ke, Re, Ue = oracle.element_nl(Xe, u[idx], dat[e])
        Kd[np.ix_(idx, idx)] += Ke
        Rd[idx] += Re
        Ud += Ue
    Kb = np.zeros((n, n))
    for i in range(m.nnodes):
        for p in range(rp[i], rp[i + 1]):
            j = cp[p]
            Kb[6 * i:6 * i + 6, 6 * j:6 * j + 6] = kv[36 * p:36 * p + 36].reshape(6, 6)
    assert np.abs(Kb - Kd).max() <= 1e-12 * np.abs(Kd).max()
    assert np.abs(res - Rd).max() <= 1e-12 * np.abs(Rd).max()
    assert abs(U - Ud) <= 1e-12 * abs(Ud)
