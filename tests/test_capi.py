"""CPU-side checks of libnbody_hip.so: it loads, exports every entry point the
header declares, and its host-only pieces (defaults, generators) behave."""
import math
import os
import re

import numpy as np

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "nbody_capi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(nb):
    lib = nb.load()
    names = declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(lib, name), f"libnbody_hip.so does not export {name}"
    assert sorted(nb.EXPORTS) == names


def test_default_params_are_reference_defaults(nb, oracle):
    p = nb.default_params()
    assert p.G == oracle.G_ref()  # nBodyAlgorithm.hpp:55-61, same expression
    assert p.eps2 == 1e-22
    assert p.theta == 1.05
    assert p.precision == nb.PREC_FP64
    assert p.sort_bodies == 1
    assert p.storage_size_param == 16 and p.stack_size_param == 16


def test_strerror(nb):
    assert nb.strerror(0) == "ok"
    assert "storage" in nb.strerror(nb.NB_E_NODE_OVERFLOW)


def test_star_disk_generator(nb, oracle):
    n = 4097
    a = nb.gen_star_disk(n, 123)
    b = nb.gen_star_disk(n, 123)
    c = nb.gen_star_disk(n, 124)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    assert not np.array_equal(a["x"], c["x"])
    assert a["m"][0] == 1.98847e30 and a["x"][0] == 0.0 and a["vx"][0] == 0.0
    mu = oracle.G_ref() * 1.98847e30
    r = np.sqrt(a["x"] ** 2 + a["y"] ** 2 + a["z"] ** 2)[1:]
    v2 = (a["vx"] ** 2 + a["vy"] ** 2 + a["vz"] ** 2)[1:]
    # vis-viva: semi-major axis from the specific orbital energy
    sma = -mu / (2 * (0.5 * v2 - mu / r))
    assert sma.min() >= 2.1 * (1 - 1e-9) and sma.max() <= 3.3 * (1 + 1e-9)
    hx = a["y"] * a["vz"] - a["z"] * a["vy"]
    hy = a["z"] * a["vx"] - a["x"] * a["vz"]
    hz = a["x"] * a["vy"] - a["y"] * a["vx"]
    h = np.sqrt(hx ** 2 + hy ** 2 + hz ** 2)[1:]
    ecc = np.sqrt(np.maximum(0, 1 - h ** 2 / (mu * sma)))
    assert ecc.max() <= 0.3 + 1e-6
    inc = np.degrees(np.arccos(hz[1:] / h))
    assert inc.max() <= 20 + 1e-6
    msun = a["m"][1:] / 1.98847e30
    assert msun.min() >= 1e-20 * (1 - 1e-12) and msun.max() <= 1e-12 * (1 + 1e-12)
    # log-uniform masses: mean of log10 near -16
    assert abs(np.log10(msun).mean() + 16) < 0.2


def test_plummer_generator(nb, oracle):
    n = 3000
    p = nb.gen_plummer(n, 7)
    G = oracle.G_ref()
    assert math.isclose(G * p["m"].sum(), 1.0, rel_tol=1e-12)
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        assert abs(p[k].mean()) < 1e-12
    r = np.sqrt(p["x"] ** 2 + p["y"] ** 2 + p["z"] ** 2)
    assert r.max() < 101
    # half-mass radius of a Plummer sphere with a = 1 is 1/sqrt(2^(2/3)-1) = 1.305
    assert abs(np.median(r) - 1.305) < 0.1
    # virial equilibrium 2K/|U| ~ 1 (G = 1 units, masses 1/n)
    m = 1.0 / n
    K = 0.5 * m * (p["vx"] ** 2 + p["vy"] ** 2 + p["vz"] ** 2).sum()
    pos = np.stack([p["x"], p["y"], p["z"]], 1)
    U = 0.0
    for i in range(0, n, 500):
        d = np.linalg.norm(pos[i:i + 500, None, :] - pos[None, :, :], axis=2)
        idx = np.arange(i, min(i + 500, n))
        mask = np.arange(n)[None, :] > idx[:, None]
        U -= (m * m / d[mask]).sum()
    assert abs(2 * K / abs(U) - 1) < 0.1
