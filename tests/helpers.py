"""Shared input makers and checks for the test suite."""
import math

import numpy as np


def uniform_cube(n, seed, lo=-1.0, hi=1.0, mass=(1e20, 1e24)):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(lo, hi, n) for _ in range(3))
    m = rng.uniform(mass[0], mass[1], n)
    v = [rng.normal(0, 1e-3, n) for _ in range(3)]
    return dict(x=x, y=y, z=z, vx=v[0], vy=v[1], vz=v[2], m=m)


def clustered(n, seed):
    """Gaussian clumps: deep, uneven trees."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-5, 5, (8, 3))
    which = rng.integers(0, 8, n)
    scale = 10.0 ** rng.uniform(-4, -1, 8)
    p = centers[which] + rng.normal(0, 1, (n, 3)) * scale[which][:, None]
    m = rng.uniform(1e22, 1e24, n)
    v = rng.normal(0, 1e-3, (n, 3))
    return dict(x=p[:, 0].copy(), y=p[:, 1].copy(), z=p[:, 2].copy(), vx=v[:, 0].copy(),
                vy=v[:, 1].copy(), vz=v[:, 2].copy(), m=m)


def digit(px, py, pz, mn, e):
    """Octant digit with the reference's FP64 chain (Subtrees.cpp:348-358)."""
    up = py > mn[1] + (e / 2.0)
    right = px > mn[0] + (e / 2.0)
    back = pz < mn[2] + (e / 2.0)
    return int(up) * 4 + int(right) * 2 + int(back)


def child_min(mn, e, d):
    """Child cell minimum (Subtrees.cpp:256-315): shift x on bit 2, y on bit 4,
    z when bit 1 is clear."""
    ce = e / 2
    return (mn[0] + ce if d & 2 else mn[0], mn[1] + ce if d & 4 else mn[1],
            mn[2] if d & 1 else mn[2] + ce)


def brute_force_canonical(x, y, z, m, aabb):
    """Recursive partition into the octree where a cell is split iff it holds
    two or more bodies, emitted depth-first in digit order.  Independent of the
    insertion-based builders; pins their topology, COM sums and counts."""
    n = len(x)
    rec = dict(depth=[], kind=[], body=[], mass=[], com_x=[], com_y=[], com_z=[], count=[])

    def emit(depth, kind, body, mass, cx, cy, cz, cnt):
        rec["depth"].append(depth)
        rec["kind"].append(kind)
        rec["body"].append(body)
        rec["mass"].append(mass)
        rec["com_x"].append(cx)
        rec["com_y"].append(cy)
        rec["com_z"].append(cz)
        rec["count"].append(cnt)

    def build(bodies, mn, e, depth):
        # returns (mass, cx, cy, cz, count) and emits records in pre-order
        if len(bodies) == 0:
            emit(depth, 0, n, 0.0, 0.0, 0.0, 0.0, 0)
            return (0.0, 0.0, 0.0, 0.0, 0)
        if len(bodies) == 1:
            b = bodies[0]
            v = (m[b], x[b] * m[b], y[b] * m[b], z[b] * m[b], 1)
            emit(depth, 1, b, *v)
            return v
        idx = len(rec["depth"])
        emit(depth, 2, n, 0.0, 0.0, 0.0, 0.0, 0)
        parts = [[] for _ in range(8)]
        for b in bodies:
            parts[digit(x[b], y[b], z[b], mn, e)].append(b)
        sm = sx = sy = sz = 0.0
        sc = 0
        for d in range(8):
            cm, cx, cy, cz, cc = build(parts[d], child_min(mn, e, d), e / 2, depth + 1)
            sx += cx
            sy += cy
            sz += cz
            sm += cm
            sc += cc
        rec["mass"][idx] = sm
        rec["com_x"][idx] = sx
        rec["com_y"][idx] = sy
        rec["com_z"][idx] = sz
        rec["count"][idx] = sc
        return (sm, sx, sy, sz, sc)

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10000))
    try:
        build(list(range(n)), (aabb[0], aabb[1], aabb[2]), aabb[6], 0)
    finally:
        sys.setrecursionlimit(old)
    return {k: np.array(v) for k, v in rec.items()}


def assert_canonical_equal(a, b, what=""):
    assert len(a["depth"]) == len(b["depth"]), f"{what}: node count {len(a['depth'])} != {len(b['depth'])}"
    for k in ("depth", "kind", "body", "count"):
        np.testing.assert_array_equal(np.asarray(a[k]).astype(np.int64),
                                      np.asarray(b[k]).astype(np.int64), err_msg=f"{what}: {k}")
    for k in ("mass", "com_x", "com_y", "com_z"):
        # bit-exact
        np.testing.assert_array_equal(np.asarray(a[k]).view(np.int64), np.asarray(b[k]).view(np.int64),
                                      err_msg=f"{what}: {k}")


def rel_err(ax, ay, az, rx, ry, rz):
    num = np.sqrt((ax - rx) ** 2 + (ay - ry) ** 2 + (az - rz) ** 2)
    den = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
    return num / den


def steps_for(dt, t_end):
    """Iterations of the reference loop (NaiveAlgorithm.cpp:116, :131, :257)."""
    t = 0.0
    t += dt
    k = 0
    while t <= t_end + 0.000001:
        k += 1
        t += dt
    return k


def hour():
    return 1.0 / 24


def fsum3(a):
    return math.fsum(a)
