"""Generators of the workloads (plates, cylinders, wingboxes): the
reference plate's simply supported BC pattern and point-sampled loads
(examples/plate_demo/_plate_utils.h:51-82, :146-184, :186-212) as known
answers derived by hand from the reference's loops, load resultants of the
consistent loads against closed forms, and (seed, SHA-256) fixtures of the
seeded generators' output (tests/golden/generators_sha256.json, made by
tests/golden/make_generator_fixtures.py)."""
import json
import math
import os
import sys

import numpy as np
import pytest

from gpu_fem_amd import meshes as M

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)


def test_reference_ss_pattern_known_answer():
    # _plate_utils.h:51-82 on a 2 x 2 element plate (3 x 3 nodes), in the
    # loops' order: corner node 0 (DOFs 0-5); iy = 0 edge, ix = 1..2: v, w;
    # for iy = 1..2: x = 0 node u, w, then x = Lx node w; iy = 2 edge,
    # ix = 1..nnx-2: w
    expected = [0, 1, 2, 3, 4, 5,      # node 0
                7, 8, 13, 14,          # nodes 1, 2: dof 2 (v), 3 (w)
                18, 20, 32,            # iy = 1: node 3 (u, w), node 5 (w)
                36, 38, 50,            # iy = 2: node 6 (u, w), node 8 (w)
                44]                    # node 7 (w)
    assert M.plate_bcs_ss_reference(2, 2).tolist() == expected
    # count: 6 + 2 (nnx - 1) + 3 (nny - 1) + (nnx - 2); 2,405 at 400 x 400
    assert len(M.plate_bcs_ss_reference(400, 400)) == 2405
    assert len(M.plate_bcs_ss_reference(7, 3)) == 6 + 2 * 7 + 3 * 3 + 6
    m = M.plate(5, 4, bc="ss", load="sine")
    assert np.array_equal(m.bcs, M.plate_bcs_ss_reference(5, 4))
    assert np.all(m.loads[m.bcs] == 0.0)


def test_reference_point_sampled_load_known_answer():
    # _plate_utils.h:146-184: w = load sin(PI x / Lx) sin(PI y / Ly) at the
    # nodes with the literal PI = 3.1415926535897 (so the x = Lx edge is not
    # exactly zero: sin(3.1415926535897) = 9.3e-14)
    PI = 3.1415926535897
    f = M.plate_loads_nodal_reference(2, 2, 1.0, 2.0, 1000.0)
    w = f[2::6].reshape(3, 3)  # [iy, ix]
    assert w[1, 1] == pytest.approx(1000.0 * math.sin(PI * 0.5) *
                                    math.sin(PI * 1.0 / 2.0), rel=1e-15)
    assert w[1, 2] == pytest.approx(1000.0 * math.sin(PI) *
                                    math.sin(PI * 0.5), rel=1e-12)
    assert 0.0 < w[1, 2] < 1e-9
    assert np.all(w[0] == 0.0)  # y = 0 row: sin(0)
    mask = np.ones(f.size, bool)
    mask[2::6] = False
    assert np.all(f[mask] == 0.0)


def test_consistent_plate_loads_resultants():
    # uniform p over Lx x Ly; sine load integrates to p 4 Lx Ly / pi^2
    m = M.plate(12, 8, 2.0, 1.0, bc="ss", load="uniform", load_mag=700.0)
    raw = M.consistent_pressure_loads(m.conn, m.xpts,
                                      lambda x: np.full(len(x), 700.0))
    assert raw[2::6].sum() == pytest.approx(700.0 * 2.0, rel=1e-13)
    s = M.consistent_pressure_loads(
        m.conn, m.xpts,
        lambda x: 1e3 * np.sin(np.pi * x[:, 0] / 2.0) * np.sin(np.pi * x[:, 1]))
    assert s[2::6].sum() == pytest.approx(1e3 * 4 * 2.0 / np.pi ** 2, rel=1e-4)


def test_cylinder_loads_and_bcs():
    nc, na, R, L, p, q = 32, 12, 1.0, 3.0, 50e3, 10e3
    m = M.cylinder(nc, na, R=R, L=L, pressure=p, axial_line_load=q, seed=4)
    # clamped end: all six DOFs of the z = 0 ring
    ring0 = np.flatnonzero(np.abs(m.xpts[:, 2]) < 1e-12)
    assert sorted(m.bcs.tolist()) == sorted((6 * ring0[:, None] +
                                             np.arange(6)).ravel().tolist())
    raw = M.consistent_pressure_loads(
        m.conn, m.xpts, lambda x: np.full(len(x), p),
        normal_fn=lambda x: np.stack([x[:, 0], x[:, 1], 0 * x[:, 0]], 1) / R)
    fr = raw[0::6] * m.xpts[:, 0] / R + raw[1::6] * m.xpts[:, 1] / R
    # internal pressure: the radial forces add to p 2 pi R L up to the
    # faceting of the circle (flat facets, radial direction at each point:
    # O((pi / nc)^2))
    area = 2 * math.pi * R * L
    assert abs(fr.sum() - p * area) <= 2 * (math.pi / nc) ** 2 * p * area
    # a closed ring under uniform pressure: no net force
    assert abs(raw[0::6].sum()) < 1e-9 * p * area
    assert abs(raw[1::6].sum()) < 1e-9 * p * area
    # axial line load on the z = L ring: q times the circumference
    ringL = np.flatnonzero(np.abs(m.xpts[:, 2] - L) < 1e-12)
    assert m.loads[6 * ringL + 2].sum() == pytest.approx(q * 2 * math.pi * R,
                                                         rel=1e-12)
    # thickness per element uniform in the range, from the seed
    t = m.data[:, 2]
    assert t.min() >= 1e-3 and t.max() <= 8e-3
    assert np.array_equal(t, M.cylinder(nc, na, seed=4).data[:, 2])
    assert not np.array_equal(t, M.cylinder(nc, na, seed=5).data[:, 2])


def test_served_requests_ranges():
    reqs = M.served_requests(40, seed=3)
    for m in reqs:
        t = m.data[0, 2]
        assert 1e-3 <= t <= 10e-3 and np.all(m.data[:, 2] == t)
        if m.name.startswith("plate"):
            n = m.meta["nxe"]
            assert 32 <= n <= 256 and m.meta["nye"] == n
        else:
            nc, na = m.meta["nc"], m.meta["na"]
            assert 64 <= nc <= 192 and 128 <= na <= 384
    kinds = {m.name[:5] for m in reqs}
    assert kinds == {"plate", "cylin"}


def _digest():
    import make_generator_fixtures as F

    return F


def test_seeded_generators_match_fixtures():
    F = _digest()
    gold = json.load(open(os.path.join(GOLD, "generators_sha256.json")))
    for name, mk in F.cases().items():
        assert F.digest(mk()) == gold[name], name
    got = [dict(F.digest(m), name=m.name) for m in F.served_cases()]
    assert got == gold["served_requests_8_seed0"]
