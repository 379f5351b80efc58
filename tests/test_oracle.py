"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Known answers come from the reference's own tests (tests/BarnesHutTest.cpp);
the rest from closed forms, an independent recursive octree builder and
long-double brute force.
"""
import math

import numpy as np
import pytest

from helpers import (assert_canonical_equal, brute_force_canonical, clustered, rel_err, steps_for,
                     uniform_cube)

X3 = [0.0, 0.0, 2.0]
Y3 = [1.0, 0.0, 0.0]
Z3 = [0.0, 2.0, 0.0]
M3 = [10.0, 10.0, 10.0]


# ---- reference fixtures (tests/BarnesHutTest.cpp) -------------------------------
def test_aabb_fixture(oracle):
    # BarnesHutTest.cpp:11-33
    a = oracle.aabb(X3, Y3, Z3)
    assert a[6] == 2
    assert (a[0], a[1], a[2]) == (0, -0.5, 0)
    assert (a[3], a[4], a[5]) == (2, 1.5, 2)


def test_synchronized_topology_fixture(oracle):
    # BarnesHutTest.cpp:35-73
    t = oracle.Tree(3, 16 * 3)
    t.build(X3, Y3, Z3, M3, builder=oracle.BUILDER_SYNCHRONIZED)
    e = t.export()
    assert list(e["body_of_node"][:9]) == [3, 0, 3, 3, 3, 3, 2, 1, 3]


def test_subtrees_topology_fixture(oracle):
    # BarnesHutTest.cpp:75-127 (maxBuildLevel 1, storage 3*3)
    t = oracle.Tree(3, 3 * 3)
    t.build(X3, Y3, Z3, M3, builder=oracle.BUILDER_SUBTREES, max_build_level=1, top_wi=3,
            subtree_wi=3)
    e = t.export()
    assert list(e["body_of_node"][:9]) == [3, 0, 3, 3, 3, 3, 2, 1, 3]
    assert e["sum_masses"][0] == 30.0
    assert list(e["sorted"]) == [1, 2, 0]


def test_fixture_digits(oracle):
    # the fixture's slots translated to octant digits: slot s holds digit
    # {5,7,4,6,1,3,0,2}[s]; body0 -> digit 5, body2 -> digit 3, body1 -> digit 0
    t = oracle.Tree(3, 48)
    t.build(X3, Y3, Z3, M3)
    e = t.export()
    oct_ = e["octants"]
    assert e["body_of_node"][oct_[5][0]] == 0
    assert e["body_of_node"][oct_[3][0]] == 2
    assert e["body_of_node"][oct_[0][0]] == 1


# ---- AABB closed forms ---------------------------------------------------------
def test_aabb_includes_origin(oracle):
    # partial extrema start at 0.0 (BarnesHutOctree.cpp:58-72)
    a = oracle.aabb([1.0, 2.0], [1.5, 1.75], [1.25, 1.5])
    assert (a[0], a[1], a[2]) == (0.0, 0.0, 0.0) or a[0] == 0.0
    assert a[3] == 2.0 and a[6] == 2.0


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_aabb_cubify(oracle, axis):
    rng = np.random.default_rng(axis)
    p = rng.uniform(-1, 1, (3, 500))
    p[axis] *= 5
    a = oracle.aabb(*p)
    lo = np.minimum(p.min(axis=1), 0.0)
    hi = np.maximum(p.max(axis=1), 0.0)
    ext = hi - lo
    e = ext.max()
    assert a[6] == e
    for c in range(3):
        if c == int(np.argmax(ext)):
            assert a[c] == lo[c] and a[3 + c] == hi[c]
        else:
            assert a[c] == lo[c] - (e - ext[c]) / 2
            assert a[3 + c] == hi[c] + (e - ext[c]) / 2


# ---- topology vs an independent recursive builder -------------------------------
@pytest.mark.parametrize("maker,n,seed", [(uniform_cube, 1, 0), (uniform_cube, 2, 1),
                                          (uniform_cube, 257, 2), (clustered, 400, 3),
                                          (uniform_cube, 1500, 4)])
@pytest.mark.parametrize("builder,level", [("subtrees", 7), ("subtrees", 1), ("subtrees", 3),
                                           ("synchronized", 7)])
def test_canonical_vs_bruteforce(oracle, maker, n, seed, builder, level):
    d = maker(n, seed)
    t = oracle.Tree(n)
    b = oracle.BUILDER_SUBTREES if builder == "subtrees" else oracle.BUILDER_SYNCHRONIZED
    t.build(d["x"], d["y"], d["z"], d["m"], builder=b, max_build_level=level)
    ref = brute_force_canonical(list(d["x"]), list(d["y"]), list(d["z"]), list(d["m"]), t.aabb)
    assert_canonical_equal(t.canonical(), ref, f"{builder}/{level}/n={n}")


def test_tree_invariants(oracle):
    d = clustered(5000, 11)
    t = oracle.Tree(5000)
    t.build(d["x"], d["y"], d["z"], d["m"])
    c = t.canonical()
    internal = int((c["kind"] == 2).sum())
    assert t.node_count == 1 + 8 * internal
    leaves = c["body"][c["kind"] == 1]
    assert sorted(leaves.tolist()) == list(range(5000))
    # in-order sort == order of body leaves in the digit-ordered DFS (R10)
    e = t.export()
    np.testing.assert_array_equal(e["sorted"], leaves)
    assert c["count"][0] == 5000
    # root mass == sum of masses (closed form, up to summation rounding)
    assert math.isclose(c["mass"][0], math.fsum(d["m"]), rel_tol=1e-13)
    assert math.isclose(c["com_x"][0], math.fsum(d["x"] * d["m"]), rel_tol=1e-9, abs_tol=1e10)


def test_builders_agree(oracle):
    d = clustered(3000, 5)
    t1 = oracle.Tree(3000)
    t1.build(d["x"], d["y"], d["z"], d["m"], builder=oracle.BUILDER_SUBTREES)
    t2 = oracle.Tree(3000)
    t2.build(d["x"], d["y"], d["z"], d["m"], builder=oracle.BUILDER_SYNCHRONIZED)
    assert_canonical_equal(t1.canonical(), t2.canonical(), "subtrees vs synchronized")


# ---- naive forces ---------------------------------------------------------------
def test_constants(oracle):
    G = oracle.G_ref()
    assert math.isclose(G, 6.67428e-11 * (86400.0 ** 2) / (1.49597870691e11 ** 3), rel_tol=1e-14)
    assert oracle.eps2_ref() == 1e-22


def test_naive_two_body(oracle):
    G = oracle.G_ref()
    x, y, z, m = [0.0, 1.5], [0.0, -0.5], [0.0, 0.25], [2e30, 3e24]
    ax, ay, az = oracle.naive_accel(x, y, z, m)
    r = np.array([1.5, -0.5, 0.25])
    r2 = float(r @ r) + 1e-22
    expect = G * m[1] * r / r2 ** 1.5
    np.testing.assert_allclose([ax[0], ay[0], az[0]], expect, rtol=1e-15)
    expect1 = -G * m[0] * r / r2 ** 1.5
    np.testing.assert_allclose([ax[1], ay[1], az[1]], expect1, rtol=1e-15)


def test_naive_single_body_is_zero(oracle):
    ax, ay, az = oracle.naive_accel([0.3], [0.1], [-2.0], [5e20])
    assert ax[0] == 0.0 and ay[0] == 0.0 and az[0] == 0.0


def test_naive_ring_symmetry(oracle):
    k = 12
    ang = 2 * np.pi * np.arange(k) / k
    x = np.concatenate([[0.0], np.cos(ang)])
    y = np.concatenate([[0.0], np.sin(ang)])
    z = np.zeros(k + 1)
    m = np.concatenate([[1e30], np.full(k, 1e25)])
    ax, ay, az = oracle.naive_accel(x, y, z, m)
    single = oracle.G_ref() * 1e25
    assert math.hypot(ax[0], ay[0]) < 1e-12 * single
    mag = np.hypot(ax[1:], ay[1:])
    np.testing.assert_allclose(mag, mag[0], rtol=1e-12)
    # every ring body is pulled towards the centre
    dots = ax[1:] * x[1:] + ay[1:] * y[1:]
    assert np.all(dots < 0)


def test_naive_newton_third_law(oracle):
    d = uniform_cube(800, 7)
    ax, ay, az = oracle.naive_accel(d["x"], d["y"], d["z"], d["m"])
    scale = np.sum(d["m"] * np.sqrt(ax ** 2 + ay ** 2 + az ** 2))
    for a in (ax, ay, az):
        assert abs(math.fsum(d["m"] * a)) < 1e-12 * scale


def test_naive_vs_longdouble(oracle):
    d = uniform_cube(96, 8)
    G = oracle.G_ref()
    ax, ay, az = oracle.naive_accel(d["x"], d["y"], d["z"], d["m"])
    L = np.longdouble
    x, y, z, m = (np.asarray(d[k], dtype=L) for k in ("x", "y", "z", "m"))
    rx = x[None, :] - x[:, None]
    ry = y[None, :] - y[:, None]
    rz = z[None, :] - z[:, None]
    r2 = rx * rx + ry * ry + rz * rz + L(1e-22)
    w = m[None, :] / (r2 * np.sqrt(r2))
    ex = (w * rx).sum(axis=1) * L(G)
    ey = (w * ry).sum(axis=1) * L(G)
    ez = (w * rz).sum(axis=1) * L(G)
    err = rel_err(ax, ay, az, ex.astype(float), ey.astype(float), ez.astype(float))
    assert err.max() < 1e-13


def test_naive_targets_match_full(oracle):
    d = uniform_cube(500, 9)
    full = oracle.naive_accel(d["x"], d["y"], d["z"], d["m"])
    t = np.array([0, 17, 499, 250])
    sub = oracle.naive_accel_targets(d["x"], d["y"], d["z"], d["m"], t)
    for a, b in zip(full, sub):
        np.testing.assert_array_equal(a[t], b)


# ---- leapfrog --------------------------------------------------------------------
def test_leapfrog_closed_form(oracle):
    rng = np.random.default_rng(3)
    n = 100
    x = [rng.normal(size=n) for _ in range(3)]
    v = [rng.normal(size=n) for _ in range(3)]
    a = [rng.normal(size=n) for _ in range(3)]
    dt = 0.125  # power of two: every product below is exact
    x0 = [c.copy() for c in x]
    vh = oracle.kick_drift(dt, a, v, x)
    for c in range(3):
        np.testing.assert_array_equal(vh[c], v[c] + a[c] * (dt / 2.0))
        np.testing.assert_allclose(x[c], x0[c] + v[c] * dt + 0.5 * a[c] * dt * dt, rtol=1e-14, atol=1e-15)
    v2 = oracle.kick(dt, a, vh)
    for c in range(3):
        np.testing.assert_allclose(v2[c], v[c] + a[c] * dt, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("dt,t_end,steps", [(1 / 24, 24 / 24, 24), (1 / 24, 10 / 24, 10),
                                            (0.001, 0.005, 5)])
def test_loop_step_count(oracle, dt, t_end, steps):
    assert steps_for(dt, t_end) == steps
    d = uniform_cube(40, 1)
    r = oracle.simulate("naive", d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"], dt, t_end)
    assert r["steps"] == steps


def test_two_body_kepler_orbit(oracle):
    # circular orbit of a light body around 1 Msun at 1 AU: after one period
    # the leapfrog returns close to the start and conserves energy to O(dt^2)
    G = oracle.G_ref()
    M = 1.98847e30
    mu = G * M
    v = math.sqrt(mu / 1.0)
    period = 2 * math.pi / math.sqrt(mu)
    nsteps = 4000
    dt = period / nsteps
    x, y, z = [0.0, 1.0], [0.0, 0.0], [0.0, 0.0]
    vx, vy, vz = [0.0, 0.0], [0.0, v], [0.0, 0.0]
    m = [M, 1.0]
    r = oracle.simulate("naive", x, y, z, vx, vy, vz, m, dt, dt * nsteps)
    assert r["steps"] == steps_for(dt, dt * nsteps) == nsteps
    pos = np.array([r["x"][1] - r["x"][0], r["y"][1] - r["y"][0]])
    assert np.linalg.norm(pos - np.array([1.0, 0.0])) < 1e-5
    e0 = 0.5 * v * v - mu / 1.0
    vv = np.array([r["vx"][1], r["vy"][1]])
    e1 = 0.5 * vv @ vv - mu / np.linalg.norm(pos)
    assert abs(e1 - e0) < 1e-6 * abs(e0)


# ---- Barnes-Hut traversal ---------------------------------------------------------
def test_bh_theta_zero_is_direct_sum(oracle):
    # theta = 0 accepts only body leaves: the direct sum in traversal order
    d = clustered(700, 12)
    t = oracle.Tree(700)
    t.build(d["x"], d["y"], d["z"], d["m"])
    ax, ay, az, inter = t.bh_accel(0.0)
    assert np.all(inter == 699)
    rx, ry, rz = oracle.naive_accel(d["x"], d["y"], d["z"], d["m"])
    # Differences come only from the summation order and from the leaf COM
    # (x*m)/m, which may move a source by an ulp: bound both.
    px, py, pz = (np.asarray(d[k]) for k in ("x", "y", "z"))
    dx = px[None, :] - px[:, None]
    dy = py[None, :] - py[:, None]
    dz = pz[None, :] - pz[:, None]
    r2 = dx * dx + dy * dy + dz * dz + 1e-22
    G = oracle.G_ref()
    absum = G * (d["m"][None, :] * np.sqrt(r2 - 1e-22) / r2 ** 1.5).sum(axis=1)
    ulp = 2.0 ** -51 * np.maximum(np.maximum(abs(px), abs(py)), abs(pz))
    moved = 3 * G * (d["m"][None, :] * ulp[None, :] / r2 ** 1.5).sum(axis=1)
    diff = np.sqrt((ax - rx) ** 2 + (ay - ry) ** 2 + (az - rz) ** 2)
    assert np.all(diff <= 1e-13 * absum + moved)


def test_bh_error_falls_with_theta(oracle):
    d = uniform_cube(4000, 13)
    t = oracle.Tree(4000)
    t.build(d["x"], d["y"], d["z"], d["m"])
    rx, ry, rz = oracle.naive_accel(d["x"], d["y"], d["z"], d["m"])
    errs = []
    for th in (0.3, 0.6, 1.0):
        ax, ay, az, _ = t.bh_accel(th)
        errs.append(np.median(rel_err(ax, ay, az, rx, ry, rz)))
    assert errs[0] < errs[1] < errs[2]
    assert errs[0] < 1e-3


def test_bh_two_body_exact(oracle):
    x, y, z, m = [0.1, 0.9], [0.2, 0.4], [0.3, 0.3], [1e24, 2e24]
    t = oracle.Tree(2)
    t.build(x, y, z, m)
    ax, ay, az, _ = t.bh_accel(1.0)
    rx, ry, rz = oracle.naive_accel(x, y, z, m)
    assert rel_err(ax, ay, az, rx, ry, rz).max() < 1e-15


def test_bh_sort_flag_does_not_change_forces(oracle):
    d = uniform_cube(600, 14)
    t = oracle.Tree(600)
    t.build(d["x"], d["y"], d["z"], d["m"])
    a1 = t.bh_accel(0.7, sort_bodies=True)
    a2 = t.bh_accel(0.7, sort_bodies=False)
    for u, v in zip(a1[:3], a2[:3]):
        np.testing.assert_array_equal(u, v)


# ---- failure handling (reference: silent corruption or endless loops) -------------
def test_duplicate_bodies_overflow_storage(oracle):
    t = oracle.Tree(3)
    with pytest.raises(oracle.OracleError) as e:
        t.build([0.5, 0.5, -0.2], [0.5, 0.5, 0.1], [0.5, 0.5, 0.3], [1.0, 1.0, 1.0])
    assert e.value.code == oracle.ORC_E_NODE_OVERFLOW


def test_zero_mass_reported(oracle):
    t = oracle.Tree(3)
    with pytest.raises(oracle.OracleError) as e:
        t.build([0.5, -0.5, 0.1], [0.5, 0.2, 0.1], [0.5, 0.1, -0.3], [1.0, 0.0, 1.0])
    assert e.value.code == oracle.ORC_E_ZERO_MASS


def test_tiny_stack_reported(oracle):
    d = uniform_cube(2000, 15)
    t = oracle.Tree(2000)
    t.build(d["x"], d["y"], d["z"], d["m"])
    with pytest.raises(oracle.OracleError) as e:
        t.bh_accel(0.2, stack=12)
    assert e.value.code == oracle.ORC_E_STACK_OVERFLOW


def test_stack_size_rule(oracle):
    # Configuration.cpp:24-33
    assert oracle.stack_size(16384, 16) == 16 * 14
    assert oracle.stack_size(1000, 16) == 16 * 10 + 500
    assert oracle.stack_size(1 << 20, 16) == 320


# ---- energy ------------------------------------------------------------------------
def test_energy_two_body_closed_form(oracle):
    G = oracle.G_ref()
    M, m, r = 1.98847e30, 5.97e24, 1.0
    v = math.sqrt(G * M / r)
    e = oracle.energy([0.0, r], [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [0.0, v], [0.0, 0.0], [M, m])
    assert math.isclose(e["kinetic"], 0.5 * m * v * v, rel_tol=1e-15)
    assert math.isclose(e["potential"], -G * M * m / r, rel_tol=1e-15)
    assert math.isclose(e["total"], e["kinetic"] + e["potential"], rel_tol=1e-15)
    # circular orbit: 2K = |U| for the orbiting body
    assert math.isclose(e["virial"], 1.0, rel_tol=1e-12)


def test_energy_matches_numpy(oracle):
    d = uniform_cube(300, 21)
    e = oracle.energy(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
    G = oracle.G_ref()
    p = np.stack([d["x"], d["y"], d["z"]], 1)
    dist = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
    iu = np.triu_indices(300, 1)
    U = -(G * d["m"][iu[0]] * d["m"][iu[1]] / dist[iu]).sum()
    K = 0.5 * (d["m"] * (d["vx"] ** 2 + d["vy"] ** 2 + d["vz"] ** 2)).sum()
    assert math.isclose(e["potential"], U, rel_tol=1e-12)
    assert math.isclose(e["kinetic"], K, rel_tol=1e-12)
