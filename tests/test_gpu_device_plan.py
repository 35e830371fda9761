"""The plan built on the device (fmm_plan_build_device) against the host plan
(fmm_plan_build, whose tree and lists equal the oracle's).

Tree: the same cells, cell numbering, centres, radii and body permutation,
bit for bit (every body's octant path comes from the same centre arithmetic).
Lists: the device traversal runs the dual walk breadth first, so a target's
list comes out ordered by source cell while the host list keeps the visiting
order: the lists must hold the same pairs, i.e. every target's sorted host
list equals its device list, with equal pair and body-pair counts; and a
solve on either plan gives the same potentials and accelerations (to the
rounding that the summation order leaves, 1e-12).  Cases: uniform cube and
Plummer sphere, theta 0.3-0.6, ncrit 8-128, nspawn 100 and 5000 (the
both-sides split rule), more than ncrit coincident bodies (a leaf at the
depth cap), one body, no bodies, and the 2^20-body Plummer and 2^21-body
cube at the benchmark parameters.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _F():
    import exafmm_amd
    return exafmm_amd


def _same_tree(hp, dp):
    ht, dt = hp.tree(), dp.tree()
    assert ht["depth"] == dt["depth"]
    for k in ("perm", "iparent", "ichild", "nchild", "ibody", "nbody", "X", "R"):
        assert ht[k].shape == dt[k].shape and np.array_equal(ht[k], dt[k]), k


def _same_lists(hp, dp):
    _same_tree(hp, dp)
    hs, ds = hp.stats(), dp.stats()
    for k in ("ncell", "m2l_pairs", "p2p_pairs", "p2p_body_pairs"):
        assert hs[k] == ds[k], (k, hs[k], ds[k])
    hm_ptr, hm_src, hp_ptr, hp_src = hp.lists()
    dm_ptr, dm_src, dp_ptr, dp_src = dp.lists()
    assert np.array_equal(hm_ptr, dm_ptr) and np.array_equal(hp_ptr, dp_ptr)
    for ptr, hsrc, dsrc in ((hm_ptr, hm_src, dm_src), (hp_ptr, hp_src, dp_src)):
        # every target's sources, sorted: sort by (target, source) in one go
        tgt = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
        order = np.lexsort((hsrc, tgt))
        assert np.array_equal(hsrc[order], dsrc)
        # the device list is sorted within each target
        assert np.all((np.diff(dsrc) > 0) | (np.diff(tgt) != 0))
    return hs


@pytest.mark.parametrize("dist,n,theta,ncrit,nspawn", [
    ("cube1", 20000, 0.4, 64, 5000),
    ("cube1", 20000, 0.6, 16, 100),
    ("plummer1", 30000, 0.4, 32, 5000),
    ("plummer1", 30000, 0.3, 8, 1000),
    ("cube1", 5000, 0.5, 128, 5000),
])
def test_device_plan_equals_host_plan(dist, n, theta, ncrit, nspawn):
    F = _F()
    x, q = F.make_bodies(n, dist, 3)
    f = F.FMM(P=6, theta=theta, ncrit=ncrit, nspawn=nspawn)
    hp = F.Plan(x, ncrit=ncrit, theta=theta, nspawn=nspawn)
    dp = F.Plan(x, ncrit=ncrit, theta=theta, nspawn=nspawn, lists_on=f)
    s = _same_lists(hp, dp)
    assert s["m2l_pairs"] > 0 and s["p2p_pairs"] > 0
    f.set_bodies_plan(hp, q)
    f.evaluate()
    th = f.targets()
    f.set_bodies_plan(dp, q)
    f.evaluate()
    td = f.targets()
    f.close()
    rel = np.linalg.norm(td - th, axis=0) / np.linalg.norm(th, axis=0)
    assert np.all(rel < 1e-12), rel


def test_device_plan_edge_cases():
    F = _F()
    f = F.FMM(P=4)
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, (3000, 3))
    x[:300] = x[0]  # 300 coincident bodies: one leaf at the depth cap
    for xs in (x, x[:1], np.zeros((0, 3))):
        hp = F.Plan(xs, ncrit=64)
        dp = F.Plan(xs, ncrit=64, lists_on=f)
        _same_lists(hp, dp)
    f.close()


@pytest.mark.parametrize("dist,n", [("plummer1", 1 << 20), ("cube1", 1 << 21)])
def test_device_plan_full_size(dist, n):
    F = _F()
    x, q = F.make_bodies(n, dist, 0)
    f = F.FMM(P=4)
    hp = F.Plan(x, ncrit=64, theta=0.4, nspawn=5000)
    dp = F.Plan(x, ncrit=64, theta=0.4, nspawn=5000, lists_on=f)
    s = _same_lists(hp, dp)
    print("%s n=%d: %d M2L, %d P2P pairs; tree host %.1f ms, device %.1f ms; lists host %.1f ms, device %.1f ms"
          % (dist, n, s["m2l_pairs"], s["p2p_pairs"], hp.stats()["host_tree_ms"], dp.stats()["host_tree_ms"],
             hp.stats()["host_lists_ms"], dp.stats()["host_lists_ms"]))
    f.close()


@pytest.mark.parametrize("P,prec,dist", [(10, "fp64", "plummer1"), (12, "fp64", "cube1"), (8, "mixed", "plummer1"),
                                         (12, "mixed", "cube1"), (12, "mixed16", "plummer1"), (4, "fp64", "cube1")])
def test_device_groups_and_tiles_solve_like_host(P, prec, dist):
    """M2L groups and tiles built with the plan on the device: the same
    groups (count) and the same solve as the host-built ones, to the rounding
    of the order in which a tile adds its slices (1e-12)."""
    F = _F()
    x, q = F.make_bodies(40000, dist, 4)
    pr = {"fp64": F.FMM_FP64, "mixed": F.FMM_MIXED, "mixed16": F.FMM_MIXED16}[prec]
    f = F.FMM(P=P, precision=pr)
    hp = F.Plan(x)
    dp = F.Plan(x, lists_on=f)
    sd = dp.stats()
    assert sd["m2l_groups"] > 0  # built on the device
    # the local expansions after the upward pass and M2L alone: the P2P of
    # the mixed path sums each target's sources in FP32 runs, whose values
    # follow the list order (source order on the device, visiting order on
    # the host), so the whole solve is compared at FP64 only
    out = {}
    for name, plan in (("host", hp), ("device", dp)):
        f.set_bodies_plan(plan, q)
        f.upward()
        f.m2l()
        f.synchronize()
        out[name] = (f.expansions()[1], f.stats())
        if prec == "fp64":
            f.evaluate()
            out[name] += (f.targets(),)
    f.close()
    (Lh, sh), (Ld, st) = out["host"][:2], out["device"][:2]
    assert st["m2l_path"] == F.M2L_TILES and sh["m2l_path"] == F.M2L_TILES
    assert st["m2l_groups"] == sh["m2l_groups"] == sd["m2l_groups"] and st["m2l_ungrouped"] == 0
    assert np.linalg.norm(Ld - Lh) / np.linalg.norm(Lh) < 1e-12
    if prec == "fp64":
        th, td = out["host"][2], out["device"][2]
        rel = np.linalg.norm(td - th, axis=0) / np.linalg.norm(th, axis=0)
        assert np.all(rel < 1e-12), rel


def _same_pairs(h, d):
    """Lists (ptr, src, ptr, src) holding the same pairs, the device ones
    sorted by source within each target."""
    for hptr, hsrc, dptr, dsrc in ((h[0], h[1], d[0], d[1]), (h[2], h[3], d[2], d[3])):
        assert np.array_equal(hptr, dptr)
        tgt = np.repeat(np.arange(len(hptr) - 1), np.diff(hptr))
        assert np.array_equal(hsrc[np.lexsort((hsrc, tgt))], dsrc)


def test_device_plan_arrays_held_on_the_device():
    """A device-built plan keeps its lists' source columns and its tiles'
    pair arrays on the device, and a context on that device copies them
    device to device.  Checked: the context's lists (read back from its
    device copies) hold the host plan's pairs; the plan can go right after
    the upload; a context of another kind (mixed precision, P = 8, from a
    plan grouped for FP64 P = 10) groups from the lists fetched to the host
    and matches the host plan after M2L (1e-12); a merge of device plans
    (host copies fetched) solves like the merge of the host plans (1e-12)."""
    F = _F()
    x, q = F.make_bodies(30000, "plummer1", 5)
    x2, q2 = F.make_bodies(12000, "cube1", 6)
    f = F.FMM(P=10)
    hp = F.Plan(x)
    dp = F.Plan(x, lists_on=f)
    assert dp.stats()["m2l_groups"] > 0
    f.set_bodies_plan(dp, q)
    del dp
    _same_pairs(hp.lists(), f.lists())
    f.evaluate()
    td = f.targets()
    f.set_bodies_plan(hp, q)
    f.evaluate()
    th = f.targets()
    rel = np.linalg.norm(td - th, axis=0) / np.linalg.norm(th, axis=0)
    assert np.all(rel < 1e-12), rel
    # another kind of context: host grouping from the fetched lists
    g = F.FMM(P=8, precision=F.FMM_MIXED)
    dp = F.Plan(x, lists_on=f)
    L = {}
    for name, plan in (("host", hp), ("device", dp)):
        g.set_bodies_plan(plan, q)
        g.upward()
        g.m2l()
        g.synchronize()
        L[name] = g.expansions()[1]
    assert np.linalg.norm(L["device"] - L["host"]) / np.linalg.norm(L["host"]) < 1e-12
    g.close()
    # merged device plans against merged host plans
    dp2, hp2 = F.Plan(x2, lists_on=f), F.Plan(x2)
    md, mh = F.Plan.merge([dp, dp2]), F.Plan.merge([hp, hp2])
    qq = np.concatenate([q, q2])
    out = []
    for m in (mh, md):
        f.set_bodies_plan(m, qq)
        f.evaluate()
        out.append(f.targets())
    f.close()
    rel = np.linalg.norm(out[1] - out[0], axis=0) / np.linalg.norm(out[0], axis=0)
    assert np.all(rel < 1e-12), rel
    _same_pairs(mh.lists(), md.lists())


def test_set_bodies_device():
    """fmm_set_bodies_device: the device plan uploaded without a copy.  The
    same solve as fmm_plan_build_device + fmm_set_bodies_plan (1e-14: split
    tiles add their parts into L atomically) and as fmm_set_bodies (1e-12);
    the timings of the plan are in the context's stats; periodic contexts
    are refused."""
    F = _F()
    x, q = F.make_bodies(30000, "plummer1", 7)
    f = F.FMM(P=10)
    out = []
    f.set_bodies_device(x, q)
    s = f.stats()
    for load in (lambda: None, lambda: f.set_bodies_plan(F.Plan(x, lists_on=f), q), lambda: f.set_bodies(x, q)):
        load()
        f.evaluate()
        out.append(f.targets())
    f.close()
    assert s["m2l_groups"] > 0 and s["host_tree_ms"] > 0 and s["plan_groups_ms"] > 0
    for b, tol in ((out[1], 1e-14), (out[2], 1e-12)):
        rel = np.linalg.norm(out[0] - b, axis=0) / np.linalg.norm(b, axis=0)
        assert np.all(rel < tol), rel
    g = F.FMM(P=6, images=1, cycle=2 * np.pi)
    with pytest.raises(F.FMMError):
        g.set_bodies_device(x, q)
    g.close()
