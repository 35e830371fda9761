"""Mixed precision (FMM_MIXED: FP32 P2P and FP32-MFMA M2L, FP64 elsewhere)
against the FP64 path and the FP64 direct sum.

Tolerances: the FP32 list phases must stay well below the FMM truncation
error they sit beside -- rel-L2 distance to the FP64 FMM below 1e-6 for the
potential and 5e-6 for the acceleration, and an error against the FP64 direct
sum within 1e-6 (pot) / 5e-6 (acc) of the FP64 FMM's own error.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

POT_TOL, ACC_TOL = 1e-6, 5e-6


def _F():
    import exafmm_amd
    return exafmm_amd


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _solve(F, x, q, P, precision, **kw):
    f = F.FMM(P=P, precision=precision, **kw)
    f.set_bodies(x, q)
    f.evaluate()
    out = f.targets(input_order=True), f.stats()
    f.close()
    return out


@pytest.mark.parametrize("P,dist,n,mode", [(4, "cube1", 20000, 0), (8, "cube1", 40000, 0),
                                           (10, "plummer1", 40000, 0), (12, "cube1", 40000, 0),
                                           (8, "cube1", 40000, 2), (12, "plummer1", 30000, 2),
                                           (10, "cube1", 40000, 3), (12, "plummer1", 30000, 3)])
def test_mixed_matches_fp64(P, dist, n, mode):
    F = _F()
    x, q = F.make_bodies(n, dist, 1)
    t64, s64 = _solve(F, x, q, P, F.FMM_FP64)
    t32, s32 = _solve(F, x, q, P, F.FMM_MIXED, m2l_mode=mode)
    # AUTO runs mixed M2L in target tiles at every order
    assert s32["m2l_path"] == (mode if mode != F.M2L_AUTO else F.M2L_TILES)
    assert s32["m2l_groups"] > 0 and s32["m2l_ungrouped"] == 0
    dp, da = _rel(t32[:, 0], t64[:, 0]), _rel(t32[:, 1:], t64[:, 1:])
    idx = np.arange(1000) * (n // 1000)
    d = O.direct(O.Kernel(P), x[idx], x, q)
    e64, e32 = O.rel_l2(d, t64[idx]), O.rel_l2(d, t32[idx])
    print("P=%d %s n=%d: mixed-fp64 %.2e %.2e; vs direct fp64 %.2e %.2e mixed %.2e %.2e"
          % (P, dist, n, dp, da, e64[0], e64[1], e32[0], e32[1]))
    assert dp < POT_TOL and da < ACC_TOL
    assert e32[0] < e64[0] + POT_TOL and e32[1] < e64[1] + ACC_TOL


def test_mixed_single_leaf_and_empty():
    F = _F()
    x, q = F.make_bodies(50, "cube1", 2)
    t32, s = _solve(F, x, q, 8, F.FMM_MIXED)
    assert s["m2l_pairs"] == 0
    d = O.direct(O.Kernel(8), x, x, q)
    assert _rel(t32[:, 0], d[:, 0]) < POT_TOL and _rel(t32[:, 1:], d[:, 1:]) < ACC_TOL
    t0, _ = _solve(F, np.zeros((0, 3)), np.zeros(0), 8, F.FMM_MIXED)
    assert t0.shape == (0, 4)


@pytest.mark.parametrize("mode", [0, 2])
def test_mixed_caller_tree_with_uneven_radii(mode):
    """A caller-built tree whose cells on one level do not share a radius
    (half the cells of every level 0.1% larger): the FP32 operators are scaled
    by the radii of the pairs they were built for, so a pair whose radii differ
    from its operator's must run per target in FP64.  Against the FP64 solve
    of the same tree and lists, at the tolerances above."""
    F = _F()
    x, q = F.make_bodies(20000, "cube1", 3)
    t = O.Tree(x, 64)
    R = t.R.copy()
    R[1::2] *= 1.001
    tree = {"iparent": t.iparent, "ichild": t.ichild, "nchild": t.nchild, "ibody": t.ibody,
            "nbody": t.nbody, "X": t.X, "R": R}
    out = []
    for prec in (F.FMM_FP64, F.FMM_MIXED):
        f = F.FMM(P=8, precision=prec, m2l_mode=mode)
        f.set_tree(x[t.perm], q[t.perm], tree)
        f.evaluate()
        out.append((f.targets(input_order=False), f.stats()))
        f.close()
    (t64, s64), (t32, s32) = out
    assert s32["m2l_pairs"] == s64["m2l_pairs"] and s32["m2l_ungrouped"] > 0
    dp, da = _rel(t32[:, 0], t64[:, 0]), _rel(t32[:, 1:], t64[:, 1:])
    print("uneven radii, mode %d: ungrouped %d of %d; mixed-fp64 %.2e %.2e"
          % (mode, s32["m2l_ungrouped"], s32["m2l_pairs"], dp, da))
    assert dp < POT_TOL and da < ACC_TOL


@pytest.mark.parametrize("case", ["ncrit512", "coincident300"])
def test_mixed_leaves_above_256_bodies(case):
    """Leaves of more than 64 lanes x 4 targets = 256 bodies (ncrit above 256,
    or more than ncrit coincident bodies kept in one leaf at the depth cap)
    take the P2P kernels' loop over target chunks."""
    F = _F()
    n = 30000
    x, q = F.make_bodies(n, "cube1", 4)
    ncrit, theta = 64, 0.4
    if case == "ncrit512":
        ncrit = 512
    else:
        n = 6000
        x, q = x[:n].copy(), q[:n] - q[:n].mean()
        x[1000:1300] = x[7]
        theta = 1e-6  # every interaction P2P: the solve is the direct sum
    t64, _ = _solve(F, x, q, 8, F.FMM_FP64, ncrit=ncrit, theta=theta)
    t32, s = _solve(F, x, q, 8, F.FMM_MIXED, ncrit=ncrit, theta=theta)
    idx = np.concatenate([np.arange(1000, 1300), np.arange(0, n, 97)])
    k = O.Kernel(8)
    d = O.direct(k, x[idx], x, q)
    e64, e32 = O.rel_l2(d, t64[idx]), O.rel_l2(d, t32[idx])
    print("%s: leaves %d; vs direct fp64 %.2e %.2e mixed %.2e %.2e" % (case, s["nleaf"], *e64, *e32))
    assert e64[0] < (1e-4 if theta > 0.1 else 1e-13) and e64[1] < (1e-4 if theta > 0.1 else 1e-13)
    assert e32[0] < e64[0] + POT_TOL and e32[1] < e64[1] + ACC_TOL
