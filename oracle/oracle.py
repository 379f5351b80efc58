"""ctypes view of oracle/liboracle.so -- the CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg.  The product never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)

ORC_OK = 0
ORC_E_INVALID_ARG = 1
ORC_E_NODE_OVERFLOW = 3
ORC_E_STACK_OVERFLOW = 4
ORC_E_ZERO_MASS = 6
ORC_E_AABB = 8

BUILDER_SUBTREES = 0
BUILDER_SYNCHRONIZED = 1


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} failed with oracle code {code}")
        self.code = code


def _load():
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError(f"{_LIB_PATH} is not built; run `make -C oracle`")
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_G_ref.restype = ctypes.c_double
    lib.orc_eps2_ref.restype = ctypes.c_double
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_get_threads.restype = ctypes.c_int
    lib.orc_naive_accel.argtypes = [ctypes.c_uint32] + [_dp] * 4 + [ctypes.c_double] * 2 + [_dp] * 3
    lib.orc_naive_accel_targets.argtypes = ([ctypes.c_uint32] + [_dp] * 4 + [ctypes.c_double] * 2
                                            + [ctypes.c_uint32, _u32p] + [_dp] * 3)
    lib.orc_kick_drift.argtypes = [ctypes.c_uint32, ctypes.c_double] + [_dp] * 12
    lib.orc_kick.argtypes = [ctypes.c_uint32, ctypes.c_double] + [_dp] * 9
    lib.orc_energy.argtypes = [ctypes.c_uint32] + [_dp] * 7 + [ctypes.c_double, _dp]
    lib.orc_aabb.argtypes = [ctypes.c_uint32, ctypes.c_int] + [_dp] * 4
    lib.orc_aabb.restype = ctypes.c_int
    lib.orc_tree_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.orc_tree_create.restype = ctypes.c_void_p
    lib.orc_tree_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_tree_build.argtypes = ([ctypes.c_void_p, ctypes.c_int] + [_dp] * 4
                                   + [ctypes.c_int] * 5 + [_dp])
    lib.orc_tree_build.restype = ctypes.c_int
    lib.orc_tree_node_count.argtypes = [ctypes.c_void_p]
    lib.orc_tree_node_count.restype = ctypes.c_uint32
    lib.orc_tree_get_aabb.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_tree_export.argtypes = ([ctypes.c_void_p, _u32p, _u32p, _i32p] + [_dp] * 4
                                    + [_u32p] + [_dp] * 4 + [_u32p])
    lib.orc_tree_canonical.argtypes = ([ctypes.c_void_p, ctypes.c_uint64, _u8p, _u8p, _u32p]
                                       + [_dp] * 4 + [_u32p])
    lib.orc_tree_canonical.restype = ctypes.c_uint64
    lib.orc_bh_accel.argtypes = ([ctypes.c_void_p] + [_dp] * 3 + [ctypes.c_double] * 3
                                + [ctypes.c_uint32, ctypes.c_int] + [_dp] * 3 + [_u32p])
    lib.orc_bh_accel.restype = ctypes.c_int
    lib.orc_stack_size.argtypes = [ctypes.c_uint32, ctypes.c_int]
    lib.orc_stack_size.restype = ctypes.c_uint32
    lib.orc_simulate.argtypes = ([ctypes.c_int, ctypes.c_uint32] + [_dp] * 7
                                 + [ctypes.c_double] * 3 + [ctypes.c_int] * 3 + [_dp] * 3
                                 + [_i32p, _dp, _dp])
    lib.orc_simulate.restype = ctypes.c_int
    _fp = ctypes.POINTER(ctypes.c_float)
    lib.orc_mlp_forward.argtypes = [_fp, ctypes.c_uint32, _fp, _fp, ctypes.c_int]
    lib.orc_mlp_grad.argtypes = [_fp, ctypes.c_uint32, _fp, _fp, _dp, _dp, ctypes.c_int]
    lib.orc_mlp_adam.argtypes = ([ctypes.c_uint32, _fp, _fp, _fp, _dp] + [ctypes.c_float] * 4
                                 + [ctypes.c_uint32])
    lib.orc_mlp_pair.argtypes = ([ctypes.c_uint32] + [_dp] * 5 + [ctypes.c_double] * 4
                                 + [_dp, _fp, _dp])
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def G_ref():
    return lib().orc_G_ref()


def eps2_ref():
    return lib().orc_eps2_ref()


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def naive_accel(x, y, z, m, G=None, eps2=None):
    x, y, z, m = map(_c64, (x, y, z, m))
    n = x.size
    G = G_ref() if G is None else G
    eps2 = eps2_ref() if eps2 is None else eps2
    ax, ay, az = (np.empty(n) for _ in range(3))
    lib().orc_naive_accel(n, _d(x), _d(y), _d(z), _d(m), G, eps2, _d(ax), _d(ay), _d(az))
    return ax, ay, az


def naive_accel_targets(x, y, z, m, targets, G=None, eps2=None):
    x, y, z, m = map(_c64, (x, y, z, m))
    t = np.ascontiguousarray(targets, dtype=np.uint32)
    G = G_ref() if G is None else G
    eps2 = eps2_ref() if eps2 is None else eps2
    ax, ay, az = (np.empty(t.size) for _ in range(3))
    lib().orc_naive_accel_targets(x.size, _d(x), _d(y), _d(z), _d(m), G, eps2, t.size,
                                  t.ctypes.data_as(_u32p), _d(ax), _d(ay), _d(az))
    return ax, ay, az


def kick_drift(dt, a, v, x):
    """a, v, x: lists of 3 float64 arrays; x is updated in place; returns v_half."""
    a = [_c64(c) for c in a]
    v = [_c64(c) for c in v]
    n = a[0].size
    vh = [np.empty(n) for _ in range(3)]
    lib().orc_kick_drift(n, dt, *map(_d, a), *map(_d, v), *map(_d, vh), *map(_d, x))
    return vh


def kick(dt, a, vh):
    a = [_c64(c) for c in a]
    vh = [_c64(c) for c in vh]
    n = a[0].size
    v = [np.empty(n) for _ in range(3)]
    lib().orc_kick(n, dt, *map(_d, a), *map(_d, vh), *map(_d, v))
    return v


def energy(x, y, z, vx, vy, vz, m, G=None):
    """{kinetic, potential, total, virial} as nBodyAlgorithm::computeEnergy."""
    arrs = [_c64(a) for a in (x, y, z, vx, vy, vz, m)]
    out = np.zeros(4)
    lib().orc_energy(arrs[0].size, *map(_d, arrs), G_ref() if G is None else G, _d(out))
    return dict(zip(["kinetic", "potential", "total", "virial"], out))


def aabb(x, y, z, work_items=1024):
    x, y, z = map(_c64, (x, y, z))
    out = np.zeros(7)
    rc = lib().orc_aabb(x.size, work_items, _d(x), _d(y), _d(z), _d(out))
    if rc:
        raise OracleError(rc, "orc_aabb")
    return out


def stack_size(n, param=16):
    return lib().orc_stack_size(n, param)


class Tree:
    """Reference-layout octree built by the oracle."""

    def __init__(self, n, storage=None):
        self.n = int(n)
        self.storage = int(storage if storage is not None else 16 * n)
        self._h = lib().orc_tree_create(self.n, self.storage)
        if not self._h:
            raise OracleError(ORC_E_INVALID_ARG, "orc_tree_create")

    def close(self):
        if self._h:
            lib().orc_tree_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def build(self, x, y, z, m, builder=BUILDER_SUBTREES, max_build_level=7, aabb_wi=1024,
              top_wi=1024, subtree_wi=640, sort_bodies=True):
        self._x, self._y, self._z, self._m = map(_c64, (x, y, z, m))
        ph = np.zeros(8)
        rc = lib().orc_tree_build(self._h, builder, _d(self._x), _d(self._y), _d(self._z),
                                  _d(self._m), max_build_level, aabb_wi, top_wi, subtree_wi,
                                  int(sort_bodies), _d(ph))
        if rc:
            raise OracleError(rc, "orc_tree_build")
        return ph

    @property
    def node_count(self):
        return lib().orc_tree_node_count(self._h)

    @property
    def aabb(self):
        out = np.zeros(7)
        lib().orc_tree_get_aabb(self._h, _d(out))
        return out

    def export(self):
        S, n = self.storage, self.n
        r = dict(octants=np.zeros(8 * S, np.uint32), body_of_node=np.zeros(S, np.uint32),
                 node_is_leaf=np.zeros(S, np.int32), sum_masses=np.zeros(S),
                 com_x=np.zeros(S), com_y=np.zeros(S), com_z=np.zeros(S),
                 body_count=np.zeros(S, np.uint32), edge=np.zeros(S), min_x=np.zeros(S),
                 min_y=np.zeros(S), min_z=np.zeros(S), sorted=np.zeros(n, np.uint32))
        u = lambda a: a.ctypes.data_as(_u32p)
        lib().orc_tree_export(self._h, u(r["octants"]), u(r["body_of_node"]),
                              r["node_is_leaf"].ctypes.data_as(_i32p), _d(r["sum_masses"]),
                              _d(r["com_x"]), _d(r["com_y"]), _d(r["com_z"]), u(r["body_count"]),
                              _d(r["edge"]), _d(r["min_x"]), _d(r["min_y"]), _d(r["min_z"]),
                              u(r["sorted"]))
        r["octants"] = r["octants"].reshape(8, S)
        return r

    def canonical(self):
        cap = self.node_count
        out = dict(depth=np.zeros(cap, np.uint8), kind=np.zeros(cap, np.uint8),
                   body=np.zeros(cap, np.uint32), mass=np.zeros(cap), com_x=np.zeros(cap),
                   com_y=np.zeros(cap), com_z=np.zeros(cap), count=np.zeros(cap, np.uint32))
        k = lib().orc_tree_canonical(self._h, cap, out["depth"].ctypes.data_as(_u8p),
                                     out["kind"].ctypes.data_as(_u8p),
                                     out["body"].ctypes.data_as(_u32p), _d(out["mass"]),
                                     _d(out["com_x"]), _d(out["com_y"]), _d(out["com_z"]),
                                     out["count"].ctypes.data_as(_u32p))
        if k != cap:
            raise OracleError(-1, "orc_tree_canonical")
        return out

    def bh_accel(self, theta, G=None, eps2=None, stack=None, sort_bodies=True):
        G = G_ref() if G is None else G
        eps2 = eps2_ref() if eps2 is None else eps2
        stack = stack_size(self.n) if stack is None else stack
        n = self.n
        ax, ay, az = (np.empty(n) for _ in range(3))
        inter = np.zeros(n, np.uint32)
        rc = lib().orc_bh_accel(self._h, _d(self._x), _d(self._y), _d(self._z), theta, G, eps2,
                                stack, int(sort_bodies), _d(ax), _d(ay), _d(az),
                                inter.ctypes.data_as(_u32p))
        if rc:
            raise OracleError(rc, "orc_bh_accel")
        return ax, ay, az, inter


def simulate(algorithm, x, y, z, vx, vy, vz, m, dt, t_end, theta=1.05, storage_param=16,
             stack_param=16, max_build_level=7):
    """Returns dict(x.., v.., a.., steps, loop_ms, phase_ms) after the reference time loop."""
    st = [np.array(a, dtype=np.float64, copy=True) for a in (x, y, z, vx, vy, vz)]
    m = _c64(m)
    n = m.size
    acc = [np.empty(n) for _ in range(3)]
    steps = ctypes.c_int32(0)
    loop = ctypes.c_double(0)
    ph = np.zeros(4)
    algo = {"naive": 0, "BarnesHut": 1}[algorithm] if isinstance(algorithm, str) else algorithm
    rc = lib().orc_simulate(algo, n, *map(_d, st), _d(m), dt, t_end, theta, storage_param,
                            stack_param, max_build_level, *map(_d, acc), ctypes.byref(steps),
                            ctypes.byref(loop), _d(ph))
    if rc:
        raise OracleError(rc, "orc_simulate")
    return dict(x=st[0], y=st[1], z=st[2], vx=st[3], vy=st[4], vz=st[5], ax=acc[0], ay=acc[1],
                az=acc[2], steps=steps.value, loop_ms=loop.value, phase_ms=ph)


# ---- learned far-field correction (nbody_mlp_oracle.h) -------------------------
MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS = 13, 64, 3, 9411
_fp32p = ctypes.POINTER(ctypes.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def mlp_forward(params, X, bf16=True):
    """Y[count, 3] of the 13-64-64-64-3 SiLU network (FP64 math, BF16 operand
    rounding where the GPU rounds when bf16 is true)."""
    p, X = _f32(params), _f32(X)
    Y = np.empty((X.shape[0], MLP_OUT), np.float32)
    lib().orc_mlp_forward(p.ctypes.data_as(_fp32p), X.shape[0], X.ctypes.data_as(_fp32p),
                          Y.ctypes.data_as(_fp32p), int(bf16))
    return Y


def mlp_grad(params, X, T, bf16=True):
    """(grad[9411] float64, loss) of L = mean (y - t)^2."""
    p, X, T = _f32(params), _f32(X), _f32(T)
    g = np.empty(MLP_PARAMS)
    loss = ctypes.c_double(0)
    lib().orc_mlp_grad(p.ctypes.data_as(_fp32p), X.shape[0], X.ctypes.data_as(_fp32p),
                       T.ctypes.data_as(_fp32p), _d(g), ctypes.byref(loss), int(bf16))
    return g, loss.value


def mlp_adam(w, m, v, grad, lr, t, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in place on float32 arrays w, m, v."""
    g = _c64(grad)
    lib().orc_mlp_adam(w.size, w.ctypes.data_as(_fp32p), m.ctypes.data_as(_fp32p),
                       v.ctypes.data_as(_fp32p), _d(g), lr, beta1, beta2, eps, t)


def mlp_pair(x, y, z, m, centre, edge, m_root, G, eps2, p):
    """(features[13] float32, target[3] float64) of one (cell, target) pair,
    by brute force over the cell's bodies."""
    x, y, z, m = _c64(x), _c64(y), _c64(z), _c64(m)
    c, q = _c64(centre), _c64(p)
    f = np.empty(13, np.float32)
    t = np.empty(3)
    lib().orc_mlp_pair(m.size, _d(x), _d(y), _d(z), _d(m), _d(c), edge, m_root, G, eps2, _d(q),
                       f.ctypes.data_as(_fp32p), _d(t))
    return f, t
