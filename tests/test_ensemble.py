"""Ensembles of independent systems: one build launch, one persistent force
launch for all of them.  Barnes-Hut systems must reproduce the one-system
FP32 traversal bit for bit (same octree, same walk); direct-sum systems are
checked against the FP64 direct sum; the loop against one-system loops."""
import numpy as np
import pytest


def _ensemble(gpu, sizes, seed):
    parts, dts = [], []
    for s, n in enumerate(sizes):
        d = gpu.gen_star_disk(int(n), seed + s) if s % 2 == 0 else gpu.gen_plummer(int(n), seed + s)
        parts.append(d)
        dts.append(1.0 / 24 if s % 2 == 0 else 1e-3)
    state = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    return parts, state, np.array(dts)


def _single(gpu, d, theta):
    ctx = gpu.Context(d["m"].size, theta=theta, precision=gpu.PREC_FP32C)
    ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
    return ctx


SIZES = [3000, 1500, 20000, 700, 9000, 2, 1, 4500, 45000, 31030]


@pytest.mark.gpu
def test_ensemble_bh_forces_match_one_system(gpu):
    parts, st, dt = _ensemble(gpu, SIZES, 100)
    with gpu.Ensemble(SIZES, theta=0.6) as e:
        e.set_tuning(bh_threshold=0, waves=4096)  # every system Barnes-Hut
        assert e.uses_bh().all()
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        ax, ay, az = e.accel()
    off = np.concatenate([[0], np.cumsum(SIZES)])
    for s, d in enumerate(parts):
        ctx = _single(gpu, d, 0.6)
        ctx.bh_build()
        ctx.bh_forces()
        rx, ry, rz = ctx.accel()
        sl = slice(off[s], off[s + 1])
        assert np.array_equal(ax[sl], rx) and np.array_equal(ay[sl], ry) and \
            np.array_equal(az[sl], rz), f"system {s} (n={SIZES[s]})"


@pytest.mark.gpu
def test_ensemble_deep_trees_match_one_system(gpu):
    """Systems whose octrees need the second key word (pairs closer than
    edge / 2^21) and deep clusters (about 31 levels), beside ordinary ones:
    canonical trees and forces bit-identical to the one-system build."""
    rng = np.random.default_rng(11)
    base = rng.uniform(-1, 1, (300, 3))
    pairs = np.concatenate([base, base + rng.normal(0, 1e-9, (300, 3))])
    centres = rng.uniform(-1, 1, (3, 3))
    clumps = (centres[:, None, :] + rng.uniform(-5e-10, 5e-10, (3, 700, 3))).reshape(-1, 3)

    def system(p, m):
        n = p.shape[0]
        return dict(x=p[:, 0].copy(), y=p[:, 1].copy(), z=p[:, 2].copy(), vx=np.zeros(n),
                    vy=np.zeros(n), vz=np.zeros(n), m=m)

    parts = [system(pairs, rng.uniform(1, 2, 600)), gpu.gen_plummer(5000, 12),
             system(clumps, rng.uniform(1e20, 1e22, 2100)), gpu.gen_star_disk(3000, 13)]
    sizes = [p["m"].size for p in parts]
    st = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    dt = np.array([1e-3, 1e-3, 1e-3, 1.0 / 24])
    # about 25 single-child levels per close pair: more nodes than 16 N
    with gpu.Ensemble(sizes, theta=0.6, storage_size_param=512) as e:
        e.set_tuning(bh_threshold=0, waves=4096)
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        ax, ay, az = e.accel()
        cans = [e.tree_canonical(s) for s in range(len(sizes))]
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, d in enumerate(parts):
        with gpu.Context(sizes[s], theta=0.6, precision=gpu.PREC_FP32C,
                         storage_size_param=512) as ctx:
            ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
            ctx.bh_build()
            one = ctx.tree_canonical()
            ctx.bh_forces()
            rx, ry, rz = ctx.accel()
        for k in one:
            assert np.array_equal(cans[s][k], one[k]), f"system {s}: {k}"
        sl = slice(off[s], off[s + 1])
        assert np.array_equal(ax[sl], rx) and np.array_equal(ay[sl], ry) and \
            np.array_equal(az[sl], rz), f"system {s} (n={sizes[s]})"


@pytest.mark.gpu
def test_ensemble_large_systems_match_one_system(gpu):
    """Systems of more than 64 blocks of 1,024 bodies, whose cells crossing
    super-block ends take the centre-of-mass pass over the cross list, beside
    a small one: canonical trees and forces bit-identical to the one-system
    build."""
    sizes = [150000, 3000, 70000]
    parts, st, dt = _ensemble(gpu, sizes, 700)
    with gpu.Ensemble(sizes, theta=0.6) as e:
        e.set_tuning(bh_threshold=0, waves=4096)
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        ax, ay, az = e.accel()
        cans = [e.tree_canonical(s) for s in range(len(sizes))]
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, d in enumerate(parts):
        with _single(gpu, d, 0.6) as ctx:
            ctx.bh_build()
            one = ctx.tree_canonical()
            ctx.bh_forces()
            rx, ry, rz = ctx.accel()
        for k in one:
            assert np.array_equal(cans[s][k], one[k]), f"system {s}: {k}"
        sl = slice(off[s], off[s + 1])
        assert np.array_equal(ax[sl], rx) and np.array_equal(ay[sl], ry) and \
            np.array_equal(az[sl], rz), f"system {s} (n={sizes[s]})"


@pytest.mark.gpu
def test_ensemble_many_systems_match_one_system(gpu):
    # 600 systems: the sort that gathers each system's slots takes 10 system
    # bits (5 radix passes, its own control words); two builds in a row reuse
    # both sorts' control words
    rng = np.random.default_rng(5)
    sizes = [int(v) for v in rng.integers(1, 200, 600)]
    sizes[0], sizes[1] = 1, 2
    parts, st, dt = _ensemble(gpu, sizes, 300)
    with gpu.Ensemble(sizes, theta=0.6) as e:
        e.set_tuning(bh_threshold=0, waves=4096)  # every system Barnes-Hut
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        first = e.accel()
        e.forces()
        ax, ay, az = e.accel()
    for a, b in zip(first, (ax, ay, az)):
        assert np.array_equal(a, b)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s in list(range(0, 600, 15)) + [599]:
        ctx = _single(gpu, parts[s], 0.6)
        ctx.bh_build()
        ctx.bh_forces()
        rx, ry, rz = ctx.accel()
        ctx.close()
        sl = slice(off[s], off[s + 1])
        assert np.array_equal(ax[sl], rx) and np.array_equal(ay[sl], ry) and \
            np.array_equal(az[sl], rz), f"system {s} (n={sizes[s]})"


@pytest.mark.gpu
@pytest.mark.parametrize("waves", [1, 16, 4096])
def test_ensemble_few_waves_match_one_system(gpu, waves):
    """Each persistent wave walks many groups in turn when there are few
    waves: every walk after a wave's first must still be made by the whole
    wave (a wave that split its lanes between items left 99% of the forces
    at zero with one wave)."""
    d = gpu.gen_star_disk(3000, 100)
    ctx = _single(gpu, d, 0.6)
    ctx.bh_build()
    ctx.bh_forces()
    rx, ry, rz = ctx.accel()
    rint = ctx.interactions()
    ctx.close()
    sizes = [3000, 700]
    st = {k: np.concatenate([d[k], d[k][:700]]) for k in d}
    with gpu.Ensemble(sizes, theta=0.6) as e:
        e.set_tuning(bh_threshold=1000, waves=waves)  # one BH system, one direct
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"],
                    np.array([1.0, 1.0]))
        e.forces()
        ax, ay, az = e.accel()
        inter = e.interactions()
    assert np.array_equal(ax[:3000], rx) and np.array_equal(ay[:3000], ry) and \
        np.array_equal(az[:3000], rz)
    assert np.array_equal(inter[:3000], rint)
    assert np.all(inter[3000:] == 700) and np.all(ax[3000:] != 0)


@pytest.mark.gpu
def test_ensemble_direct_matches_fp64(gpu):
    sizes = [777, 2048, 1500, 64, 3]
    parts, st, dt = _ensemble(gpu, sizes, 200)
    with gpu.Ensemble(sizes, theta=0.5) as e:
        e.set_tuning(bh_threshold=1 << 30, waves=1024)  # every system direct
        assert not e.uses_bh().any()
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        ax, ay, az = e.accel()
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, d in enumerate(parts):
        n = d["m"].size
        ctx = gpu.Context(n)  # FP64 reference arithmetic
        ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
        ctx.forces_direct()
        rx, ry, rz = ctx.accel()
        sl = slice(off[s], off[s + 1])
        ref = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
        err = np.sqrt((ax[sl] - rx) ** 2 + (ay[sl] - ry) ** 2 + (az[sl] - rz) ** 2) / \
            np.where(ref > 0, ref, 1)
        assert np.median(err) < 1e-5 and err.max() < 1e-3, (s, np.median(err), err.max())


@pytest.mark.gpu
def test_ensemble_loop_matches_one_system_loops(gpu):
    """Three loop iterations with per-system time steps (mixed direct and
    Barnes-Hut): every Barnes-Hut system's state equals its one-system loop."""
    sizes = [4000, 600, 12000, 300]
    parts, st, dt = _ensemble(gpu, sizes, 300)
    with gpu.Ensemble(sizes, theta=0.5) as e:
        e.set_tuning(bh_threshold=1000, waves=2048)
        bh = e.uses_bh()
        assert list(bh) == [True, False, True, False]
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        e.steps(3)
        got = e.state()
        ph = e.phase_ms()
    assert ph["Acceleration Kernel Time"] > 0 and ph["Octree creation"] > 0
    off = np.concatenate([[0], np.cumsum(sizes)])
    for s, d in enumerate(parts):
        sl = slice(off[s], off[s + 1])
        if not bh[s]:
            continue
        ctx = _single(gpu, d, 0.5)
        ctx.bh_build()
        ctx.bh_forces()
        ctx.steps(gpu.ALGO_BARNES_HUT, dt[s], 3)
        ref = ctx.state()
        for a, k in zip(got, ("x", "y", "z", "vx", "vy", "vz")):
            assert np.array_equal(a[sl], ref[k]), f"system {s} {k}"


@pytest.mark.gpu
@pytest.mark.parametrize("threshold", [0, 6000])
def test_generated_ensemble_every_body_accurate(gpu, oracle, threshold):
    """A generated ensemble (sizes 2^U[10,15.5], star+disk and Plummer
    alternating), all Barnes-Hut or mixed: every body of every system gets a
    force; at sampled targets of every system, a Barnes-Hut system is at least
    as accurate (against the FP64 direct sum) as the reference's per-body
    traversal at the same theta, and a direct-sum system is within FP32 pair
    rounding."""
    theta = 0.5
    sizes, st, dt = gpu.gen_ensemble(40, 11, lo_log2=10, hi_log2=15.5)
    with gpu.Ensemble(sizes, theta=theta) as e:
        e.set_tuning(bh_threshold=threshold, waves=8192)
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        ax, ay, az = e.accel()
        bh = e.uses_bh()
        inter = e.interactions()
    assert np.all(np.sqrt(ax ** 2 + ay ** 2 + az ** 2) > 0)
    assert np.all(inter > 0)
    off = np.concatenate([[0], np.cumsum(sizes, dtype=np.int64)])
    rng = np.random.default_rng(0)
    for s in range(sizes.size):
        sl = slice(off[s], off[s + 1])
        n = int(sizes[s])
        t = np.sort(rng.choice(n, min(n, 48), replace=False)).astype(np.uint32)
        x, y, z, m = st["x"][sl], st["y"][sl], st["z"][sl], st["m"][sl]
        rx, ry, rz = oracle.naive_accel_targets(x, y, z, m, t)
        rn = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)

        def err(a, b, c):
            return np.sqrt((a[t] - rx) ** 2 + (b[t] - ry) ** 2 + (c[t] - rz) ** 2) / rn

        e_gpu = err(ax[sl], ay[sl], az[sl])
        if bh[s]:
            tree = oracle.Tree(n)
            tree.build(x, y, z, m)
            bx, by, bz, _ = tree.bh_accel(theta)
            tree.close()
            e_ref = err(bx, by, bz)
            # the group MAC opens a cell whenever any body of the group needs
            # it; FP32 pair maths adds ~1e-6
            assert np.median(e_gpu) <= np.median(e_ref) + 2e-6, (s, n, np.median(e_gpu),
                                                                  np.median(e_ref))
            assert e_gpu.max() <= 2 * e_ref.max() + 1e-5, (s, n, e_gpu.max(), e_ref.max())
        else:
            assert np.median(e_gpu) < 1e-5 and e_gpu.max() < 1e-4, (s, n, e_gpu.max())


@pytest.mark.gpu
def test_ensemble_tuning_knobs(gpu, oracle):
    """The autotuner's tile and group parameters: direct-sum tiles of 64 and
    128 targets (one or two per lane; pair offsets are taken from the tile's
    first body, so the FP32 roundings differ) both match the FP64 direct sum
    within FP32 pair rounding; every traversal group size and packing limit
    gives Barnes-Hut forces as accurate as the reference's per-body traversal;
    invalid values are rejected."""
    sizes = [3000, 500, 9000, 130]
    parts, st, dt = _ensemble(gpu, sizes, 310)
    off = np.concatenate([[0], np.cumsum(sizes)])
    with gpu.Ensemble(sizes, theta=0.5) as e:
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        got = {}
        for tile in (64, 128):
            e.set_tuning(bh_threshold=1 << 30, waves=2048, direct_tile=tile)  # all direct
            e.forces()
            got[tile] = e.accel()
        for s, d in enumerate(parts):
            sl = slice(off[s], off[s + 1])
            ctx = gpu.Context(sizes[s])  # FP64 reference arithmetic
            ctx.set_state(d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"], d["m"])
            rx, ry, rz = ctx.forces_direct()
            ctx.close()
            ref = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
            for tile, (ax, ay, az) in got.items():
                err = np.sqrt((ax[sl] - rx) ** 2 + (ay[sl] - ry) ** 2 + (az[sl] - rz) ** 2) / ref
                assert np.median(err) < 1e-5 and err.max() < 1e-3, (tile, s, err.max())
        for gmax, pmax in ((16, 64), (24, 24), (48, 192), (64, 512)):
            e.set_tuning(bh_threshold=0, waves=4096, group_max=gmax, pack_max=pmax)
            assert e.tuning["group_max"] == gmax and e.tuning["pack_max"] == pmax
            e.forces()
            ax, ay, az = e.accel()
            for s, d in enumerate(parts):
                sl = slice(off[s], off[s + 1])
                n = sizes[s]
                t = np.arange(0, n, max(1, n // 64), dtype=np.uint32)
                rx, ry, rz = oracle.naive_accel_targets(d["x"], d["y"], d["z"], d["m"], t)
                tree = oracle.Tree(n)
                tree.build(d["x"], d["y"], d["z"], d["m"])
                bx, by, bz, _ = tree.bh_accel(0.5)
                tree.close()
                rn = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
                e_gpu = np.sqrt((ax[sl][t] - rx) ** 2 + (ay[sl][t] - ry) ** 2 +
                                (az[sl][t] - rz) ** 2) / rn
                e_ref = np.sqrt((bx[t] - rx) ** 2 + (by[t] - ry) ** 2 + (bz[t] - rz) ** 2) / rn
                assert np.median(e_gpu) <= np.median(e_ref) + 2e-6, (gmax, s, np.median(e_gpu))
        for bad in (dict(group_max=0), dict(group_max=65, pack_max=256),
                    dict(group_max=32, pack_max=16), dict(direct_tile=96)):
            kw = dict(bh_threshold=0, waves=4096)
            kw.update(bad)
            with pytest.raises(gpu.NBodyError):
                e.set_tuning(**kw)


@pytest.mark.gpu
def test_ensemble_autotune_restores_state(gpu):
    sizes, st, dt = gpu.gen_ensemble(24, 7, lo_log2=10, hi_log2=13)
    with gpu.Ensemble(sizes, theta=0.5) as e:
        e.set_state(st["x"], st["y"], st["z"], st["vx"], st["vy"], st["vz"], st["m"], dt)
        e.forces()
        before = e.state()
        best, ms = e.autotune(steps=1)
        assert ms > 0
        assert best == e.tuning
        assert best["waves"] in (1024, 2048, 4096, 8192, 16384, 32768)
        assert 1 <= best["group_max"] <= 64 and best["pack_max"] >= best["group_max"]
        assert best["direct_tile"] in (64, 128)
        after = e.state()
        for a, b in zip(before, after):
            assert np.array_equal(a, b)
        e.steps(2)  # the tuned configuration runs


@pytest.mark.gpu
def test_ensemble_rejects_bad_arguments(gpu):
    with pytest.raises(gpu.NBodyError):
        gpu.Ensemble([10, 0, 5])
    with pytest.raises(gpu.NBodyError):
        gpu.Ensemble([10, 20], precision=gpu.PREC_FP64)
