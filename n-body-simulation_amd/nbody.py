"""ctypes binding of libnbody_hip.so (include/nbody_capi.h) for tests and bench.

The library is the product; this module only marshals numpy arrays across the
C ABI.  It never falls back to another implementation: if the shared library
is missing or no HIP device is present, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbody_hip.so")

NB_OK = 0
NB_E_INVALID_ARG = 1
NB_E_HIP = 2
NB_E_NODE_OVERFLOW = 3
NB_E_STACK_OVERFLOW = 4
NB_E_DEPTH = 5
NB_E_ZERO_MASS = 6
NB_E_NOT_BUILT = 7
NB_E_AABB = 8
NB_E_NO_DEVICE = 10

PREC_FP64 = 0
PREC_FP32C = 1
PREC_FP64_FAST = 2
NOALLOC_SCRATCH_BYTES = 24 * 1024 * 1024
NOALLOC_MAX_BODIES = 1 << 23
ALGO_NAIVE = 0
ALGO_BARNES_HUT = 1

PHASES = ["AABB creation", "keys and sort", "topology", "Compute center of mass",
          "Octree creation", "Acceleration Kernel Time", "Leapfrog Part 1", "Leapfrog Part 2",
          "keys", "sort"]

# every symbol include/nbody_capi.h declares
EXPORTS = [
    "nb_default_params", "nb_create", "nb_destroy", "nb_set_params", "nb_get_params",
    "nb_body_count", "nb_last_error", "nb_strerror", "nb_device_name", "nb_set_state",
    "nb_get_state", "nb_get_accel", "nb_snapshot_begin", "nb_snapshot_wait", "nb_set_positions", "nb_forces_direct",
    "nb_forces_direct_targets", "nb_bh_build", "nb_bh_forces", "nb_kick_drift", "nb_kick",
    "nb_steps", "nb_phase_ms", "nb_tree_info", "nb_tree_sorted", "nb_tree_canonical",
    "nb_bh_interactions", "nb_gen_star_disk", "nb_gen_plummer", "nb_energy",
    "nb_mlp_default_train_opts", "nb_mlp_create", "nb_mlp_destroy", "nb_mlp_last_error",
    "nb_mlp_set_params", "nb_mlp_get_params", "nb_mlp_forward", "nb_mlp_grad",
    "nb_mlp_adam_step", "nb_mlp_train", "nb_mlp_pairs", "nb_set_correction",
    "nb_tree_quadrupoles", "nb_correction_pairs", "nb_forces_direct_noalloc", "nb_quadrupole_pairs",
]

MLP_IN, MLP_OUT, MLP_PARAMS = 13, 3, 9411


class TrainOpts(ctypes.Structure):
    _fields_ = [("epochs", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("lr", ctypes.c_float),
                ("lr_final", ctypes.c_float), ("beta1", ctypes.c_float),
                ("beta2", ctypes.c_float), ("eps", ctypes.c_float), ("augment", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class Params(ctypes.Structure):
    _fields_ = [("G", ctypes.c_double), ("eps2", ctypes.c_double), ("theta", ctypes.c_double),
                ("precision", ctypes.c_uint32), ("sort_bodies", ctypes.c_uint32),
                ("storage_size_param", ctypes.c_uint32), ("stack_size_param", ctypes.c_uint32),
                ("quadrupole", ctypes.c_uint32), ("group_max", ctypes.c_uint32),
                ("pack_max", ctypes.c_uint32)]


class NBodyError(RuntimeError):
    def __init__(self, code, what, detail=""):
        super().__init__(f"{what}: code {code} ({detail})")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_fp = ctypes.POINTER(ctypes.c_float)
_lib = None


def load(path=LIB_PATH):
    """Load the shared library and declare the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    sig = {
        "nb_default_params": (None, [ctypes.POINTER(Params)]),
        "nb_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_uint32, ctypes.POINTER(Params)]),
        "nb_destroy": (ctypes.c_int, [vp]),
        "nb_set_params": (ctypes.c_int, [vp, ctypes.POINTER(Params)]),
        "nb_get_params": (ctypes.c_int, [vp, ctypes.POINTER(Params)]),
        "nb_body_count": (ctypes.c_uint32, [vp]),
        "nb_last_error": (ctypes.c_char_p, [vp]),
        "nb_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "nb_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
        "nb_set_state": (ctypes.c_int, [vp] + [_dp] * 7),
        "nb_get_state": (ctypes.c_int, [vp] + [_dp] * 6),
        "nb_get_accel": (ctypes.c_int, [vp] + [_dp] * 3),
        "nb_snapshot_begin": (ctypes.c_int, [vp]),
        "nb_snapshot_wait": (ctypes.c_int, [vp, ctypes.POINTER(_dp)]),
        "nb_set_positions": (ctypes.c_int, [vp] + [_dp] * 3),
        "nb_forces_direct": (ctypes.c_int, [vp]),
        "nb_forces_direct_noalloc": (ctypes.c_int, [ctypes.c_uint32] + [ctypes.c_void_p] * 4 +
                                     [ctypes.c_double, ctypes.c_double, ctypes.c_uint32] +
                                     [ctypes.c_void_p] * 4),
        "nb_forces_direct_targets": (ctypes.c_int, [vp, ctypes.c_uint32, _u32p] + [_dp] * 3),
        "nb_bh_build": (ctypes.c_int, [vp]),
        "nb_bh_forces": (ctypes.c_int, [vp]),
        "nb_kick_drift": (ctypes.c_int, [vp, ctypes.c_double]),
        "nb_kick": (ctypes.c_int, [vp, ctypes.c_double]),
        "nb_steps": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_uint32]),
        "nb_phase_ms": (ctypes.c_int, [vp, _dp]),
        "nb_tree_info": (ctypes.c_int, [vp, _u64p, _u32p, _u32p, _dp]),
        "nb_tree_sorted": (ctypes.c_int, [vp, _u32p]),
        "nb_tree_canonical": (ctypes.c_uint64, [vp, ctypes.c_uint64, _u8p, _u8p, _u32p]
                              + [_dp] * 4 + [_u32p]),
        "nb_bh_interactions": (ctypes.c_int, [vp, _u32p]),
        "nb_energy": (ctypes.c_int, [vp] + [_dp] * 4),
        "nb_gen_star_disk": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64] + [_dp] * 7),
        "nb_gen_plummer": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint64] + [_dp] * 7),
        "nb_mlp_default_train_opts": (None, [ctypes.POINTER(TrainOpts)]),
        "nb_mlp_create": (ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_uint64]),
        "nb_mlp_destroy": (ctypes.c_int, [vp]),
        "nb_mlp_last_error": (ctypes.c_char_p, [vp]),
        "nb_mlp_set_params": (ctypes.c_int, [vp, _fp]),
        "nb_mlp_get_params": (ctypes.c_int, [vp, _fp]),
        "nb_mlp_forward": (ctypes.c_int, [vp, ctypes.c_uint64, _fp, _fp]),
        "nb_mlp_grad": (ctypes.c_int, [vp, ctypes.c_uint32, _fp, _fp, _fp, _fp]),
        "nb_mlp_adam_step": (ctypes.c_int, [vp, _fp, ctypes.c_float, ctypes.c_uint32]),
        "nb_mlp_train": (ctypes.c_int, [vp, ctypes.c_uint64, _fp, _fp, ctypes.POINTER(TrainOpts),
                                        _fp]),
        "nb_mlp_pairs": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_uint32, _u32p,
                                        ctypes.c_uint64, _u32p, _u32p, _fp, _fp, _u64p]),
        "nb_set_correction": (ctypes.c_int, [vp, vp, ctypes.c_double]),
        "nb_tree_quadrupoles": (ctypes.c_uint64, [vp, ctypes.c_uint64, _dp]),
        "nb_correction_pairs": (ctypes.c_int, [vp, _u64p]),
        "nb_quadrupole_pairs": (ctypes.c_int, [vp, _u64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def default_params(**kw):
    p = Params()
    load().nb_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def strerror(code):
    return load().nb_strerror(code).decode()


def device_name():
    buf = ctypes.create_string_buffer(256)
    rc = load().nb_device_name(buf, 256)
    if rc:
        raise NBodyError(rc, "nb_device_name", strerror(rc))
    return buf.value.decode()


def gen_star_disk(n, seed):
    out = [np.empty(n) for _ in range(7)]
    rc = load().nb_gen_star_disk(n, seed, *map(_d, out))
    if rc:
        raise NBodyError(rc, "nb_gen_star_disk")
    return dict(zip(["x", "y", "z", "vx", "vy", "vz", "m"], out))


def gen_plummer(n, seed):
    out = [np.empty(n) for _ in range(7)]
    rc = load().nb_gen_plummer(n, seed, *map(_d, out))
    if rc:
        raise NBodyError(rc, "nb_gen_plummer")
    return dict(zip(["x", "y", "z", "vx", "vy", "vz", "m"], out))


class Context:
    """One system of n bodies on the current HIP device."""

    def __init__(self, n, params=None, **kw):
        self.lib = load()
        self.n = int(n)
        p = params if params is not None else default_params()
        for k, v in kw.items():
            setattr(p, k, v)
        h = ctypes.c_void_p()
        rc = self.lib.nb_create(ctypes.byref(h), self.n, ctypes.byref(p))
        if rc:
            raise NBodyError(rc, "nb_create", strerror(rc))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nb_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise NBodyError(rc, what, self.lib.nb_last_error(self.h).decode())

    @property
    def params(self):
        p = Params()
        self._check(self.lib.nb_get_params(self.h, ctypes.byref(p)), "nb_get_params")
        return p

    def set_params(self, **kw):
        p = self.params
        for k, v in kw.items():
            setattr(p, k, v)
        self._check(self.lib.nb_set_params(self.h, ctypes.byref(p)), "nb_set_params")

    def set_state(self, x, y, z, vx, vy, vz, m):
        arrs = [_f64(a) for a in (x, y, z, vx, vy, vz, m)]
        assert all(a.size == self.n for a in arrs)
        self._keep = arrs
        self._check(self.lib.nb_set_state(self.h, *map(_d, arrs)), "nb_set_state")

    def set_positions(self, x, y, z):
        arrs = [_f64(a) for a in (x, y, z)]
        self._check(self.lib.nb_set_positions(self.h, *map(_d, arrs)), "nb_set_positions")

    def state(self):
        out = [np.empty(self.n) for _ in range(6)]
        self._check(self.lib.nb_get_state(self.h, *map(_d, out)), "nb_get_state")
        return dict(zip(["x", "y", "z", "vx", "vy", "vz"], out))

    def accel(self):
        out = [np.empty(self.n) for _ in range(3)]
        self._check(self.lib.nb_get_accel(self.h, *map(_d, out)), "nb_get_accel")
        return tuple(out)

    def snapshot_begin(self):
        """Start the asynchronous copy-out of x, v and a (returns at once)."""
        self._check(self.lib.nb_snapshot_begin(self.h), "nb_snapshot_begin")

    def snapshot_wait(self):
        """Wait for the last snapshot; returns copies of its nine arrays."""
        ptrs = (_dp * 9)()
        self._check(self.lib.nb_snapshot_wait(self.h, ptrs), "nb_snapshot_wait")
        keys = ["x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"]
        return {k: np.ctypeslib.as_array(ptrs[i], shape=(self.n,)).copy()
                for i, k in enumerate(keys)}

    def forces_direct(self):
        self._check(self.lib.nb_forces_direct(self.h), "nb_forces_direct")
        return self.accel()

    def forces_direct_targets(self, targets):
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        out = [np.empty(t.size) for _ in range(3)]
        self._check(self.lib.nb_forces_direct_targets(self.h, t.size, t.ctypes.data_as(_u32p),
                                                      *map(_d, out)), "nb_forces_direct_targets")
        return tuple(out)

    def bh_build(self):
        self._check(self.lib.nb_bh_build(self.h), "nb_bh_build")

    def bh_forces(self):
        self._check(self.lib.nb_bh_forces(self.h), "nb_bh_forces")
        return self.accel()

    def kick_drift(self, dt):
        self._check(self.lib.nb_kick_drift(self.h, dt), "nb_kick_drift")

    def kick(self, dt):
        self._check(self.lib.nb_kick(self.h, dt), "nb_kick")

    def steps(self, algorithm, dt, count):
        algo = {"naive": ALGO_NAIVE, "BarnesHut": ALGO_BARNES_HUT}.get(algorithm, algorithm)
        self._check(self.lib.nb_steps(self.h, algo, dt, count), "nb_steps")

    def phase_ms(self):
        out = np.zeros(len(PHASES))
        self._check(self.lib.nb_phase_ms(self.h, _d(out)), "nb_phase_ms")
        return dict(zip(PHASES, out))

    def energy(self):
        out = [ctypes.c_double() for _ in range(4)]
        self._check(self.lib.nb_energy(self.h, *[ctypes.byref(o) for o in out]), "nb_energy")
        return dict(zip(["kinetic", "potential", "total", "virial"], [o.value for o in out]))

    def tree_info(self):
        nodes = ctypes.c_uint64()
        internal = ctypes.c_uint32()
        depth = ctypes.c_uint32()
        aabb = np.zeros(7)
        self._check(self.lib.nb_tree_info(self.h, ctypes.byref(nodes), ctypes.byref(internal),
                                          ctypes.byref(depth), _d(aabb)), "nb_tree_info")
        return dict(node_count=nodes.value, internal_count=internal.value,
                    max_depth=depth.value, aabb=aabb)

    def tree_sorted(self):
        out = np.empty(self.n, np.uint32)
        self._check(self.lib.nb_tree_sorted(self.h, out.ctypes.data_as(_u32p)), "nb_tree_sorted")
        return out

    def interactions(self):
        out = np.empty(self.n, np.uint32)
        self._check(self.lib.nb_bh_interactions(self.h, out.ctypes.data_as(_u32p)),
                    "nb_bh_interactions")
        return out

    def tree_canonical(self):
        cap = self.tree_info()["node_count"]
        out = dict(depth=np.zeros(cap, np.uint8), kind=np.zeros(cap, np.uint8),
                   body=np.zeros(cap, np.uint32), mass=np.zeros(cap), com_x=np.zeros(cap),
                   com_y=np.zeros(cap), com_z=np.zeros(cap), count=np.zeros(cap, np.uint32))
        k = self.lib.nb_tree_canonical(self.h, cap, out["depth"].ctypes.data_as(_u8p),
                                       out["kind"].ctypes.data_as(_u8p),
                                       out["body"].ctypes.data_as(_u32p), _d(out["mass"]),
                                       _d(out["com_x"]), _d(out["com_y"]), _d(out["com_z"]),
                                       out["count"].ctypes.data_as(_u32p))
        if k != cap:
            raise NBodyError(-1, "nb_tree_canonical", f"returned {k}, expected {cap}")
        return out

    # ---- learned far-field correction --------------------------------------------
    def tree_quadrupoles(self):
        """Quadrupole moments of every node in canonical order, shape (nodes, 6)."""
        cap = self.tree_info()["node_count"]
        q = np.zeros((cap, 6))
        k = self.lib.nb_tree_quadrupoles(self.h, cap, _d(q))
        if k != cap:
            raise NBodyError(-1, "nb_tree_quadrupoles", f"returned {k}, expected {cap}: "
                             + self.lib.nb_last_error(self.h).decode())
        return q

    def mlp_pairs(self, theta, targets, cap):
        """Training pairs of the current octree: dict(cell, body, features, residuals, found)."""
        t = np.ascontiguousarray(targets, dtype=np.uint32)
        cell = np.zeros(cap, np.uint32)
        body = np.zeros(cap, np.uint32)
        F = np.zeros((cap, MLP_IN), np.float32)
        T = np.zeros((cap, MLP_OUT), np.float32)
        found = ctypes.c_uint64()
        self._check(self.lib.nb_mlp_pairs(self.h, theta, t.size, t.ctypes.data_as(_u32p), cap,
                                          cell.ctypes.data_as(_u32p), body.ctypes.data_as(_u32p),
                                          F.ctypes.data_as(_fp), T.ctypes.data_as(_fp),
                                          ctypes.byref(found)), "nb_mlp_pairs")
        m = min(found.value, cap)
        return dict(cell=cell[:m], body=body[:m], features=F[:m], residuals=T[:m],
                    found=found.value)

    def correction_pairs(self):
        out = ctypes.c_uint64()
        self._check(self.lib.nb_correction_pairs(self.h, ctypes.byref(out)), "nb_correction_pairs")
        return out.value

    def quadrupole_pairs(self):
        """Cell interactions evaluated with the quadrupole term in the last traversal."""
        out = ctypes.c_uint64()
        self._check(self.lib.nb_quadrupole_pairs(self.h, ctypes.byref(out)), "nb_quadrupole_pairs")
        return out.value

    def set_correction(self, mlp, tau=0.0):
        """Attach a trained Mlp (None detaches) to the FP32C Barnes-Hut traversal."""
        self._corr = mlp  # the network must outlive the attachment
        self._check(self.lib.nb_set_correction(self.h, mlp.h if mlp is not None else None, tau),
                    "nb_set_correction")


def default_train_opts(**kw):
    o = TrainOpts()
    load().nb_mlp_default_train_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Mlp:
    """The correction network (weights on the current HIP device)."""

    def __init__(self, seed=1):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.nb_mlp_create(ctypes.byref(h), seed)
        if rc:
            raise NBodyError(rc, "nb_mlp_create", strerror(rc))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nb_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise NBodyError(rc, what, self.lib.nb_mlp_last_error(self.h).decode())

    @property
    def params(self):
        p = np.empty(MLP_PARAMS, np.float32)
        self._check(self.lib.nb_mlp_get_params(self.h, p.ctypes.data_as(_fp)), "nb_mlp_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == MLP_PARAMS
        self._check(self.lib.nb_mlp_set_params(self.h, p.ctypes.data_as(_fp)), "nb_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], MLP_OUT), np.float32)
        self._check(self.lib.nb_mlp_forward(self.h, X.shape[0], X.ctypes.data_as(_fp),
                                            Y.ctypes.data_as(_fp)), "nb_mlp_forward")
        return Y

    def grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.nb_mlp_grad(self.h, X.shape[0], X.ctypes.data_as(_fp),
                                         T.ctypes.data_as(_fp), g.ctypes.data_as(_fp),
                                         ctypes.byref(loss)), "nb_mlp_grad")
        return g, loss.value

    def adam_step(self, grad, lr, t):
        g = np.ascontiguousarray(grad, dtype=np.float32)
        self._check(self.lib.nb_mlp_adam_step(self.h, g.ctypes.data_as(_fp), lr, t),
                    "nb_mlp_adam_step")

    def train(self, X, T, **opts):
        """Train on the GPU; returns the mean batch loss of each epoch."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        o = default_train_opts(**opts)
        losses = np.zeros(o.epochs, np.float32)
        self._check(self.lib.nb_mlp_train(self.h, X.shape[0], X.ctypes.data_as(_fp),
                                          T.ctypes.data_as(_fp), ctypes.byref(o),
                                          losses.ctypes.data_as(_fp)), "nb_mlp_train")
        return losses


class EnsembleTuning(ctypes.Structure):
    _fields_ = [("bh_threshold", ctypes.c_uint32), ("waves", ctypes.c_uint32),
                ("group_max", ctypes.c_uint32), ("pack_max", ctypes.c_uint32),
                ("direct_tile", ctypes.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


ENSEMBLE_SIG = {
    "nb_ensemble_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32, _u32p,
                                          ctypes.POINTER(Params)]),
    "nb_ensemble_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "nb_ensemble_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "nb_ensemble_body_count": (ctypes.c_uint64, [ctypes.c_void_p]),
    "nb_ensemble_set_state": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 8),
    "nb_ensemble_get_state": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 6),
    "nb_ensemble_get_accel": (ctypes.c_int, [ctypes.c_void_p] + [_dp] * 3),
    "nb_ensemble_forces": (ctypes.c_int, [ctypes.c_void_p]),
    "nb_ensemble_steps": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "nb_ensemble_phase_ms": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "nb_ensemble_interactions": (ctypes.c_int, [ctypes.c_void_p, _u32p]),
    "nb_ensemble_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(EnsembleTuning)]),
    "nb_ensemble_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(EnsembleTuning)]),
    "nb_ensemble_uses_bh": (ctypes.c_int, [ctypes.c_void_p, _u8p]),
    "nb_ensemble_node_count": (ctypes.c_uint64, [ctypes.c_void_p, ctypes.c_uint32]),
    "nb_ensemble_tree_canonical": (ctypes.c_uint64, [ctypes.c_void_p, ctypes.c_uint32,
                                                     ctypes.c_uint64, _u8p, _u8p, _u32p]
                                   + [_dp] * 4 + [_u32p]),
    "nb_ensemble_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32,
                                            ctypes.POINTER(EnsembleTuning), _dp]),
}
EXPORTS += list(ENSEMBLE_SIG)


class Ensemble:
    """Many independent systems advanced together (concatenated state)."""

    def __init__(self, sizes, params=None, **kw):
        self.lib = load()
        for name, (res, args) in ENSEMBLE_SIG.items():
            fn = getattr(self.lib, name)
            fn.restype, fn.argtypes = res, args
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes, dtype=np.int64)])
        self.n = int(self.offsets[-1])
        p = params if params is not None else default_params(precision=PREC_FP32C)
        for k, v in kw.items():
            setattr(p, k, v)
        h = ctypes.c_void_p()
        rc = self.lib.nb_ensemble_create(ctypes.byref(h), self.sizes.size,
                                         self.sizes.ctypes.data_as(_u32p), ctypes.byref(p))
        if rc:
            raise NBodyError(rc, "nb_ensemble_create", strerror(rc))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.nb_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc:
            raise NBodyError(rc, what, self.lib.nb_ensemble_last_error(self.h).decode())

    def set_state(self, x, y, z, vx, vy, vz, m, dt):
        arrs = [_f64(a) for a in (x, y, z, vx, vy, vz, m)]
        dt = _f64(dt)
        assert all(a.size == self.n for a in arrs) and dt.size == self.sizes.size
        self._keep = arrs + [dt]
        self._check(self.lib.nb_ensemble_set_state(self.h, *map(_d, arrs), _d(dt)),
                    "nb_ensemble_set_state")

    def state(self):
        """[x, y, z, vx, vy, vz] of all systems (concatenated)."""
        out = [np.empty(self.n) for _ in range(6)]
        self._check(self.lib.nb_ensemble_get_state(self.h, *map(_d, out)), "nb_ensemble_get_state")
        return out

    def accel(self):
        out = [np.empty(self.n) for _ in range(3)]
        self._check(self.lib.nb_ensemble_get_accel(self.h, *map(_d, out)), "nb_ensemble_get_accel")
        return out

    def forces(self):
        self._check(self.lib.nb_ensemble_forces(self.h), "nb_ensemble_forces")

    def steps(self, count):
        self._check(self.lib.nb_ensemble_steps(self.h, count), "nb_ensemble_steps")

    def phase_ms(self):
        ms = np.zeros(len(PHASES))
        self._check(self.lib.nb_ensemble_phase_ms(self.h, _d(ms)), "nb_ensemble_phase_ms")
        return dict(zip(PHASES, ms))

    def interactions(self):
        out = np.empty(self.n, np.uint32)
        self._check(self.lib.nb_ensemble_interactions(self.h, out.ctypes.data_as(_u32p)),
                    "nb_ensemble_interactions")
        return out

    def tree_canonical(self, s):
        """Canonical dump of Barnes-Hut system s (format of Context.tree_canonical)."""
        cap = self.lib.nb_ensemble_node_count(self.h, s)
        if cap == 0:
            raise NBodyError(-1, "nb_ensemble_node_count", f"system {s} has no octree")
        out = dict(depth=np.zeros(cap, np.uint8), kind=np.zeros(cap, np.uint8),
                   body=np.zeros(cap, np.uint32), mass=np.zeros(cap), com_x=np.zeros(cap),
                   com_y=np.zeros(cap), com_z=np.zeros(cap), count=np.zeros(cap, np.uint32))
        k = self.lib.nb_ensemble_tree_canonical(
            self.h, s, cap, out["depth"].ctypes.data_as(_u8p), out["kind"].ctypes.data_as(_u8p),
            out["body"].ctypes.data_as(_u32p), _d(out["mass"]), _d(out["com_x"]),
            _d(out["com_y"]), _d(out["com_z"]), out["count"].ctypes.data_as(_u32p))
        if k != cap:
            raise NBodyError(-1, "nb_ensemble_tree_canonical", f"returned {k}, expected {cap}")
        return out

    @property
    def tuning(self):
        t = EnsembleTuning()
        self._check(self.lib.nb_ensemble_get_tuning(self.h, ctypes.byref(t)), "nb_ensemble_get_tuning")
        return t.as_dict()

    def set_tuning(self, bh_threshold, waves, group_max=64, pack_max=256, direct_tile=64):
        t = EnsembleTuning(bh_threshold, waves, group_max, pack_max, direct_tile)
        self._check(self.lib.nb_ensemble_set_tuning(self.h, ctypes.byref(t)), "nb_ensemble_set_tuning")

    def uses_bh(self):
        f = np.zeros(self.sizes.size, np.uint8)
        self._check(self.lib.nb_ensemble_uses_bh(self.h, f.ctypes.data_as(_u8p)), "nb_ensemble_uses_bh")
        return f.astype(bool)

    def autotune(self, steps=2):
        t = EnsembleTuning()
        ms = ctypes.c_double()
        self._check(self.lib.nb_ensemble_autotune(self.h, steps, ctypes.byref(t), ctypes.byref(ms)),
                    "nb_ensemble_autotune")
        return t.as_dict(), ms.value


def gen_ensemble(nsys, seed, lo_log2=10, hi_log2=16):
    """Sizes floor(2^U[lo, hi]) from a seeded generator; even systems are star +
    disk (dt = 1 h), odd ones Plummer (dt = 1e-3).  Returns (sizes, state dict,
    dt per system)."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(lo_log2, hi_log2, nsys)).astype(np.uint32)
    parts, dts = [], []
    for s, n in enumerate(sizes):
        sub = (seed ^ ((s + 1) * 0x9E3779B97F4A7C15)) & 0xFFFFFFFFFFFFFFFF
        d = gen_star_disk(int(n), sub) if s % 2 == 0 else gen_plummer(int(n), sub)
        parts.append(d)
        dts.append(1.0 / 24 if s % 2 == 0 else 1e-3)
    state = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    return sizes, state, np.array(dts)
