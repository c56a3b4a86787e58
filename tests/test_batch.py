"""Batches of triangulated structures (C5): jittered simple-cubic grain lattices traced
together in one persistent launch, each grain's candidates behind a BVH.

The batch has no reference counterpart; it is pinned by the oracle: a batch structure's
per-structure tallies equal the oracle's alias-mode trace of that structure alone over
the same path ids (the brute-force scan of findCollisionSurface, RT:1989-2096), and the
GPU's non-batched mesh paths equal the oracle's bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle as O
from ptrans import gen, geometry, post

M_TWOPI = 3.14159265358979323846 * 2.0


def _modes(nb=2, nk=4, seed=3):
    from test_alias import _table, _weights
    omega, vel, mfp, weight = _table(nb, nk, seed)
    mfp = np.clip(mfp, 80.0, None)
    prob, alias = O.build_alias(_weights(omega, vel, mfp, weight))
    return omega, vel, mfp, weight, prob, alias


def _lattices():
    return [geometry.build_geometry(gen.sc_lattice(seed=11 + j, k=2 + (j % 2), m=1 + j, p=0.3,
                                                    kind="Gamma", value=2.0))
            for j in range(3)]


def test_lattice_closed_outward_and_tiles_the_start_face():
    geo = geometry.build_geometry(gen.sc_lattice(seed=6, k=3, m=2))
    from collections import Counter
    for gi, g in enumerate(geo.grains[1:], 1):
        E = Counter()
        for f in g.surfaces:
            v = f.vertices
            for t in range(3):
                E[(v[t], v[(t + 1) % 3])] += 1
            assert np.dot(np.array(f.center) - np.array(g.generatrix), f.normal) > 0
        assert all(E[(b, a)] == 1 for (a, b) in E)
    inc = [f for g in geo.grains for f in g.surfaces if f.incident]
    area = sum(0.5 * np.linalg.norm(np.cross(np.subtract(f.vertices[1], f.vertices[0]),
                                             np.subtract(f.vertices[2], f.vertices[0]))) for f in inc)
    assert abs(area - geo.box[1] * geo.box[2]) < 1e-6 * area


def test_oracle_mesh_start_grains_are_the_first_layer():
    geo = geometry.build_geometry(gen.sc_lattice(seed=6, k=3, m=2))
    gd, _g = geo.desc()
    omega, vel, mfp, weight, prob, alias = _modes()
    md, _m = geometry.modes_desc(omega, vel, mfp, weight)
    _, _, rec = O.trace_alias(gd, md, mfp, prob, alias, 4, 0, 3000, records=True)
    assert not np.any(rec["flags"] & 2)
    first_layer = {1 + 0 + 3 * (j + 3 * l) for j in range(3) for l in range(3)}
    assert set(np.unique(rec["start_grain"]).tolist()) == first_layer


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 2])
def test_gpu_mesh_paths_bitexact_vs_oracle(gpu_ctx, precision):
    geo = geometry.build_geometry(gen.sc_lattice(seed=6, k=2, m=2))
    omega, vel, mfp, weight, prob, alias = _modes()
    gd, _g = geo.desc()
    md, _m = geometry.modes_desc(omega, vel, mfp, weight)
    _, _, orec = O.trace_alias(gd, md, mfp, prob, alias, 9, 0, 1500, records=True)
    gpu_ctx.set_seed(9)
    g = gpu_ctx.geometry(geo)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    rec = gpu_ctx.replay_alias(g, m, 0, 1500, precision=precision)
    for f in ("end_side", "flags", "events", "draws", "which_grain", "start_grain"):
        np.testing.assert_array_equal(rec[f], orec[f], err_msg=f)
    np.testing.assert_array_equal(rec["pos"].view(np.uint64), orec["pos"].view(np.uint64))


def _oracle_struct_sums(geo, mods, seed, first, n):
    omega, vel, mfp, weight, prob, alias = mods
    gd, _g = geo.desc()
    md, _m = geometry.modes_desc(omega, vel, mfp, weight)
    mp, mt = O.trace_alias(gd, md, mfp, prob, alias, seed, first, n)
    return int(mp.sum()), float((mt / mfp).sum()), int(mt.sum())


@pytest.mark.gpu
@pytest.mark.parametrize("leaf", [1, 4])
def test_gpu_batch_equals_each_structure_alone(gpu_ctx, leaf):
    from ptrans.trace import Batch
    geos = _lattices()
    mods = _modes()
    omega, vel, mfp, weight, prob, alias = mods
    npaths = [3000, 5000, 2000]
    gpu_ctx.set_seed(21)
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    b = Batch(gpu_ctx, geos, npaths, leaf_size=leaf, grid=16)
    print("batch", b.info(), flush=True)
    for prec in (0, 2):
        sums, fin, st = b.trace(m, precision=prec)
        assert st["paths"] == sum(npaths)
        assert np.all(fin > 0) and np.all(fin <= st["kernel_ms"] * 1.5 + 1.0)
        for s, geo in enumerate(geos):
            n, sx, nt = _oracle_struct_sums(geo, mods, 21, b.path_off[s], npaths[s])
            assert sums[s, 0] == n, (s, sums[s], n)
            assert abs(sums[s, 1] - sx) <= 1e-12 * max(sx, 1e-300), (s, sums[s, 1], sx)
    b.close()


@pytest.mark.gpu
def test_gpu_batch_graph_replays_equal_direct_launch(gpu_ctx):
    """The (fill, trace) graph replayed several times, with the program's own eager work
    between replays: every replay equals a direct launch with the same seed."""
    from ptrans.trace import Batch
    geos = _lattices()
    omega, vel, mfp, weight, prob, alias = _modes()
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    other = gpu_ctx.geometry(geometry.build_geometry(gen.film_holes(seed=2, nphonon=1, nangle=3)))
    b = Batch(gpu_ctx, geos, [4000, 4000, 4000], leaf_size=2, grid=16)
    for rep, seed in enumerate((5, 6, 5, 7)):
        gpu_ctx.set_seed(seed)
        direct, _, _ = b.trace(m, precision=1)
        gpu_ctx.trace_alias(other, m, 50000 * (rep + 1))          # unrelated eager work
        gpu_ctx.set_seed(seed)
        graph, fin, st = b.trace(m, precision=1, graph=True)
        np.testing.assert_array_equal(graph[:, 0], direct[:, 0])
        np.testing.assert_allclose(graph, direct, rtol=1e-12, atol=1e-14)
        assert st["paths"] == 12000
    b.close()


@pytest.mark.gpu
def test_gpu_batch_mean_specularity_gradient_matches_fd(gpu_ctx):
    """d kappa_s / d p_s for every structure (the mean specularity of its facets) by the
    likelihood-ratio scores against central differences (Delta = 1e-2, common random
    numbers: the same seeds), both averaged over 8 seeds; within 3 combined errors."""
    from ptrans.trace import Batch
    geos = _lattices()
    omega, vel, mfp, weight, prob, alias = _modes()
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    n = 200000
    b = Batch(gpu_ctx, geos, n, leaf_size=4, grid=16)
    p0 = np.array([0.25, 0.5, 0.7])
    c = 0.75 * geos[0].box[0]
    b.set_specularity(p0)
    gpu_ctx.set_seed(100)
    pilot, _, _ = b.trace(m, precision=1)
    base = pilot[:, 1] / pilot[:, 0]
    lr, fd, se = [], [], []
    for seed in range(8):
        gpu_ctx.set_seed(200 + seed)
        b.set_specularity(p0)
        s0, _, _ = b.trace(m, precision=1, baseline=base)
        lr.append(c * s0[:, 3] / s0[:, 0])
        # standard error of one launch's LR estimate from its sum of squared terms
        se.append(c * np.sqrt(np.maximum(s0[:, 4] / s0[:, 0] - (s0[:, 3] / s0[:, 0]) ** 2, 0.0)
                              / s0[:, 0]))
        b.set_specularity(p0 + 1e-2)
        sp, _, _ = b.trace(m, precision=1)
        b.set_specularity(p0 - 1e-2)
        sm, _, _ = b.trace(m, precision=1)
        fd.append(c * (sp[:, 1] / sp[:, 0] - sm[:, 1] / sm[:, 0]) / 2e-2)
    lr, fd, se = np.array(lr), np.array(fd), np.array(se)
    for s in range(3):
        # the spread over seeds agrees with the per-launch standard error (8 samples)
        ratio = lr[:, s].std(ddof=1) / se[:, s].mean()
        print("structure %d  seed spread / launch stderr %.2f" % (s, ratio), flush=True)
        assert 0.3 < ratio < 2.5, ratio
        a, ea = lr[:, s].mean(), lr[:, s].std(ddof=1) / np.sqrt(8)
        f, ef = fd[:, s].mean(), fd[:, s].std(ddof=1) / np.sqrt(8)
        print("structure %d  LR %.5f +- %.5f  FD %.5f +- %.5f" % (s, a, ea, f, ef), flush=True)
        assert abs(a - f) <= 3 * np.hypot(ea, ef)
    b.close()


@pytest.mark.gpu
def test_gpu_batch_diffusive_paths_equal_oracle(gpu_ctx):
    """Mean free paths of 10 - 40 nm: the accelerated batch trace skips BVH scans inside
    safety balls (found through the grain's BVH); per-structure tallies still equal the
    oracle's brute-force trace of each structure alone."""
    from test_alias import _weights
    from ptrans.anphon import synthetic_table
    from ptrans.trace import Batch
    geos = _lattices()
    t = synthetic_table(2, 4, 9, mfp_min_nm=10.0, mfp_max_nm=40.0)
    omega = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    mfp = t.mfp_nm
    prob, alias = O.build_alias(_weights(omega, t.vel, mfp, t.weight))
    mods = (omega, t.vel, mfp, t.weight, prob, alias)
    npaths = [1500, 2500, 1000]
    gpu_ctx.set_seed(33)
    m = gpu_ctx.modes(omega=omega, vel=t.vel, mfp=mfp, weight=t.weight)
    b = Batch(gpu_ctx, geos, npaths, leaf_size=2, grid=16)
    sums, fin, st = b.trace(m, precision=2)
    print("events per path", st["events"] / st["paths"], flush=True)
    for s, geo in enumerate(geos):
        n, sx, nt = _oracle_struct_sums(geo, mods, 33, b.path_off[s], npaths[s])
        assert sums[s, 0] == n
        assert abs(sums[s, 1] - sx) <= 1e-12 * max(sx, 1e-300), (s, sums[s, 1], sx)
    b.close()


@pytest.mark.gpu
def test_gpu_batch_at_c5_sizes_equals_brute_force_oracle(gpu_ctx):
    """Structures of the screening batch's own generator (4 x 4 x 4 jittered grains) at
    1,920, 12,000 and 94,080 triangles -- the low end, middle and high end of its
    2e3 - 1e5 range, 160 - 2,352 BVH-held candidates per grain -- traced together, 2,000
    paths each: per-structure path counts and sum x equal the oracle's brute-force scan of
    each structure alone (findCollisionSurface's candidate set and first-index tie-break,
    RT:1989-2096), in FP64 and in the bit-identical FP32 filter, at leaf sizes 2 and 8."""
    from ptrans.trace import Batch
    geos = [gen.sc_lattice_flat(seed=400 + m, k=4, m=m, jitter=5.0, trans_kind=2,
                                trans_param=2.0, p=0.3) for m in (2, 5, 14)]
    mods = _modes()
    omega, vel, mfp, weight, prob, alias = mods
    npaths = [2000, 2000, 2000]
    ref = []
    for s, geo in enumerate(geos):
        ref.append(_oracle_struct_sums(geo, mods, 41, 2000 * s, 2000))
    m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
    gpu_ctx.set_seed(41)
    for leaf in (2, 8):
        b = Batch(gpu_ctx, geos, npaths, leaf_size=leaf, grid=32)
        print("C5-size batch, leaf %d:" % leaf, b.info(), flush=True)
        for prec in (0, 2):
            sums, fin, st = b.trace(m, precision=prec)
            assert st["paths"] == 6000
            for s in range(3):
                n, sx, nt = ref[s]
                assert sums[s, 0] == n, (leaf, prec, s, sums[s, 0], n)
                assert abs(sums[s, 1] - sx) <= 1e-12 * max(sx, 1e-300), (leaf, prec, s, sums[s, 1], sx)
        b.close()


_TOP_REF = {}        # oracle sums of the LDS-levels cases, shared by the parameters


@pytest.mark.gpu
@pytest.mark.parametrize("levels", [1, 2, 3])
def test_gpu_batch_lds_top_levels_equal_brute_force_oracle(gpu_ctx, levels):
    """Top BVH levels staged in LDS (restaged as the waves move from structure to
    structure): C5-size structures (1,920 / 12,000 / 94,080 triangles) and twelve small
    lattices of 400 paths each (many restagings), ballistic-leaning and diffusive (safety
    balls found through the BVH) mode tables, in FP64, the FP32 filter and the FP32 trace:
    per-structure path counts and sum x equal the oracle's brute-force scan of each
    structure alone."""
    from test_alias import _weights
    from ptrans.anphon import synthetic_table
    from ptrans.trace import Batch
    t = synthetic_table(2, 4, 9, mfp_min_nm=10.0, mfp_max_nm=40.0)
    omega_d = t.omega_cm() * 0.03 * M_TWOPI * 1.0e12
    prob_d, alias_d = O.build_alias(_weights(omega_d, t.vel, t.mfp_nm, t.weight))
    tables = {"ballistic": _modes(),
              "diffusive": (omega_d, t.vel, t.mfp_nm, t.weight, prob_d, alias_d)}
    c5 = [gen.sc_lattice_flat(seed=400 + m, k=4, m=m, jitter=5.0, trans_kind=2,
                              trans_param=2.0, p=0.3) for m in (2, 5, 14)]
    cases = [("c5", c5, [1500, 1500, 1500], ("ballistic",)),
             ("small", _lattices() * 4, [400] * 12, ("ballistic", "diffusive"))]
    for name, geos, npaths, tnames in cases:
        b = Batch(gpu_ctx, geos, npaths, leaf_size=2, grid=32)
        b.set_top_levels(levels)
        for tname in tnames:
            mods = tables[tname]
            omega, vel, mfp, weight, prob, alias = mods
            m = gpu_ctx.modes(omega=omega, vel=vel, mfp=mfp, weight=weight)
            key = (name, tname)
            if key not in _TOP_REF:
                _TOP_REF[key] = [_oracle_struct_sums(geo, mods, 43, b.path_off[s], npaths[s])
                                 for s, geo in enumerate(geos)]
            ref = _TOP_REF[key]
            gpu_ctx.set_seed(43)
            for prec in (0, 2):
                print("levels %d, %s, %s modes, precision %d" % (levels, name, tname, prec),
                      flush=True)
                sums, fin, st = b.trace(m, precision=prec, graph=prec == 2)
                assert st["paths"] == sum(npaths)
                for s in range(len(geos)):
                    n, sx, nt = ref[s]
                    assert sums[s, 0] == n, (name, tname, prec, s, sums[s, 0], n)
                    assert abs(sums[s, 1] - sx) <= 1e-12 * max(sx, 1e-300), (name, tname, prec, s)
            # the FP32 trace (not bit-identical to FP64) gives the same paths with the image
            # as without it
            b.set_top_levels(0)
            plain, _, _ = b.trace(m, precision=1)
            b.set_top_levels(levels)
            staged, _, _ = b.trace(m, precision=1, graph=True)
            np.testing.assert_array_equal(staged[:, 0], plain[:, 0])
            np.testing.assert_allclose(staged[:, :3], plain[:, :3], rtol=1e-12, atol=1e-300)
            m.close()
        b.close()
