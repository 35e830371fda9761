"""Pins the oracle (oracle/oracle.c) before it is trusted as the checker.

Each operator is checked against exafmm-beta's own compiled operators
(oracle/_ref, when the reference tree was available at build time), against
golden numbers the reference printed (tree/list counts and rel-L2 errors
recorded from the reference binary), and against what the mathematics fixes
(closed forms, translation exactness, convergence in P, linearity).
"""
import math

import numpy as np
import pytest

import oracle as O

REF = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _cluster(rng, c, r, n):
    return c + rng.uniform(-r, r, (n, 3))


# ---------------------------------------------------------------- operators vs reference
@REF
@pytest.mark.parametrize("P", [3, 4, 10, 13])
def test_operators_match_reference(P):
    rng = np.random.default_rng(P)
    ko, kr = O.Kernel(P), O.Kernel(P, use_ref=True)
    Xc = rng.uniform(-1, 1, 3)
    xb = _cluster(rng, Xc, 0.5, 37)
    q = rng.uniform(-1, 1, 37)
    Mo, Mr = ko.p2m(Xc, xb, q), kr.p2m(Xc, xb, q)
    assert _rel(Mo, Mr) < 1e-14
    Xp = Xc + np.array([0.5, -0.5, 0.5])
    Mpo, Mpr = ko.m2m(Xp, Xc, Mo), kr.m2m(Xp, Xc, Mr)
    assert _rel(Mpo, Mpr) < 1e-13
    Xi = Xp + np.array([4.0, 1.0, -3.0])
    Lo, Lr = ko.m2l(Xi, Xp, Mpo), kr.m2l(Xi, Xp, Mpr)
    assert _rel(Lo, Lr) < 1e-13
    Xl = Xi + np.array([-0.25, 0.5, 0.25])
    Llo, Llr = ko.l2l(Xl, Xi, Lo), kr.l2l(Xl, Xi, Lr)
    assert _rel(Llo, Llr) < 1e-13
    xt = _cluster(rng, Xl, 0.25, 21)
    assert _rel(ko.l2p(Xl, Llo, xt), kr.l2p(Xl, Llr, xt)) < 1e-13
    # P2P: the reference's SIMD path seeds rsqrt in float + 2 Newton steps
    assert _rel(ko.p2p(xt, xb, q), kr.p2p(xt, xb, q)) < 1e-12


@REF
def test_generators_match_reference():
    for dist, gen in (("cube", O.data_cube),):
        x, q = O.ref_init_bodies(3000, dist, 5)
        assert np.array_equal(gen(3000, 5), x)
        assert np.array_equal(O.data_source(3000, 5), q)


@pytest.mark.skipif(not O.ref_ieee_available(), reason="oracle/_ref IEEE build absent")
@pytest.mark.parametrize("n,seed", [(3000, 5), (20000, 0)])
def test_plummer_matches_reference_up_to_fast_math(n, seed):
    """orc_data_plummer (include/dataset.h:109-136) is bit-equal to the
    reference's generator compiled without -ffast-math.  The shipped build
    (-ffast-math) rounds the pow/sqrt/cos chain differently in the last bits
    (at most a few 1e-13 relative); the octree and the interaction lists built
    from either set are identical, so every count and error downstream is too."""
    x = O.data_plummer(n, seed)
    xi, qi = O.ref_init_bodies(n, "plummer", seed, ieee=True)
    assert np.array_equal(x, xi) and np.array_equal(O.data_source(n, seed), qi)
    xf, _ = O.ref_init_bodies(n, "plummer", seed)
    rel = np.abs(xf - x) / np.maximum(np.abs(x), 1e-300)
    assert rel.max() < 1e-12
    t, tf = O.Tree(x, 64), O.Tree(xf, 64)
    assert t.ncell == tf.ncell and np.array_equal(t.perm, tf.perm)
    for k in ("iparent", "ichild", "nchild", "ibody", "nbody"):
        assert np.array_equal(getattr(t, k), getattr(tf, k))
    a, b = t.traverse(0.4, 5000), tf.traverse(0.4, 5000)
    assert all(np.array_equal(u, v) for u, v in zip(a, b))


# ---------------------------------------------------------------- golden numbers from the reference binary
def test_golden_counts_cube32768():
    """examples/fmm.cxx defaults (cube, N=32768, ncrit=64, theta=.4): the reference
    builds 2,413 cells (2,111 leaves, depth 4) and visits 312,856 M2L and
    142,813 P2P pairs (41,663,490 body pairs)."""
    x = O.data_cube(32768, 0)
    t = O.Tree(x, 64)
    assert (t.ncell, int((t.nchild == 0).sum()), t.depth) == (2413, 2111, 4)
    ci, cj, ty = t.traverse(0.4, 5000)
    assert int((ty == 0).sum()) == 312856 and int((ty == 1).sum()) == 142813
    p2p = ty == 1
    assert int((t.nbody[ci[p2p]].astype(np.int64) * t.nbody[cj[p2p]]).sum()) == 41663490


def test_golden_counts_plummer1m():
    """Reference Plummer generator, N=2^20: 64,513 cells, 56,446 leaves, depth 12,
    11,883,385 M2L and 4,288,618 P2P pairs, 1,999,873,168 body pairs."""
    x = O.data_plummer(1 << 20, 0)
    t = O.Tree(x, 64)
    assert (t.ncell, int((t.nchild == 0).sum()), t.depth) == (64513, 56446, 12)
    ci, cj, ty = t.traverse(0.4, 5000)
    assert int((ty == 0).sum()) == 11883385 and int((ty == 1).sum()) == 4288618
    p2p = ty == 1
    assert int((t.nbody[ci[p2p]].astype(np.int64) * t.nbody[cj[p2p]]).sum()) == 1999873168


def test_golden_error_cube32768_p10():
    """The reference prints rel-L2 4.7429586e-06 (pot) / 4.8780842e-06 (acc) for
    the examples/fmm.cxx defaults at P=10 (100 sampled tree-order targets)."""
    n = 32768
    x, q = O.data_cube(n, 0), O.data_source(n, 0)
    t = O.Tree(x, 64)
    xs, qs = x[t.perm], q[t.perm]
    k = O.Kernel(10)
    trg = t.fmm(k, xs, qs, 0.4, 5000)
    idx = O.sample_indices(n, 100)
    pot, acc = O.rel_l2(O.direct(k, xs[idx], xs, qs), trg[idx])
    assert abs(pot - 4.7429586e-06) < 1e-13
    assert abs(acc - 4.8780842e-06) < 1e-13


@pytest.mark.parametrize("P,pot,acc", [(6, 2.29e-4, 1.14e-4), (8, 3.14e-5, 2.37e-5)])
def test_golden_error_vs_order(P, pot, acc):
    """Reference convergence table (N=32768 cube, theta=.4), 3 digits."""
    n = 32768
    x, q = O.data_cube(n, 0), O.data_source(n, 0)
    t = O.Tree(x, 64)
    xs, qs = x[t.perm], q[t.perm]
    k = O.Kernel(P)
    trg = t.fmm(k, xs, qs, 0.4, 5000)
    idx = O.sample_indices(n, 100)
    p, a = O.rel_l2(O.direct(k, xs[idx], xs, qs), trg[idx])
    assert abs(p - pot) / pot < 5e-3 and abs(a - acc) / acc < 5e-3


def _kernel_chain(P):
    """examples/kernel.cxx: one unit source through P2M -> M2M -> M2L -> L2L -> L2P."""
    k = O.Kernel(P)
    src = np.array([[2.0, 2.0, 2.0]])
    M = k.p2m([3, 1, 1], src, [1.0])
    MJ = k.m2m([4, 0, 0], [3, 1, 1], M)
    LI = k.m2l([-4, 0, 0], [4, 0, 0], MJ)
    Li = k.l2l([-3, 1, 1], [-4, 0, 0], LI)
    tgt = np.array([[-2.0, 2.0, 2.0]])
    fmm = k.l2p([-3, 1, 1], Li, tgt)
    d = k.p2p(tgt, src, [1.0])
    # kernel.cxx:93-97 normalises by the FMM value (getNrm* of `bodies`, the L2P result)
    pot = abs(fmm[0, 0] - d[0, 0]) / abs(fmm[0, 0])
    acc = np.linalg.norm(fmm[0, 1:] - d[0, 1:]) / np.linalg.norm(fmm[0, 1:])
    return pot, acc


@pytest.mark.parametrize("P,pot,acc", [(4, 5.05e-2, 1.50e-1), (10, 1.48e-3, 2.19e-2),
                                       (20, 4.65e-5, 1.06e-3)])
def test_kernel_chain_known_answer(P, pot, acc):
    """examples/kernel.cxx prints 'P potErr accErr'; the reference gives these values."""
    p, a = _kernel_chain(P)
    assert abs(p - pot) / pot < 1e-2 and abs(a - acc) / acc < 1e-2


# ---------------------------------------------------------------- mathematics
def test_p2p_closed_form_and_brute_force():
    k = O.Kernel(4)
    y = np.array([[0.3, -0.2, 0.1]])
    x = np.array([[1.0, 2.0, -0.5]])
    d = x[0] - y[0]
    r = np.linalg.norm(d)
    t = k.p2p(x, y, [2.0])[0]
    assert abs(t[0] - 2.0 / r) < 1e-15
    assert np.allclose(t[1:], -2.0 * d / r**3, rtol=1e-15, atol=0)
    rng = np.random.default_rng(0)
    xi, xj, qj = rng.normal(size=(200, 3)), rng.normal(size=(300, 3)), rng.normal(size=300)
    dx = xi[:, None, :] - xj[None, :, :]
    r = np.linalg.norm(dx, axis=2)
    pot = (qj / r).sum(1)
    grad = -(qj[None, :, None] * dx / r[:, :, None] ** 3).sum(1)
    t = k.p2p(xi, xj, qj)
    assert _rel(t[:, 0], pot) < 1e-14 and _rel(t[:, 1:], grad) < 1e-14
    # R^2 = 0 pairs are skipped
    assert np.array_equal(k.p2p(y, y, [1.0]), np.zeros((1, 4)))


def test_p2m_closed_form():
    """P2M of a unit charge on the z axis: M[n,0] = rho^n (P_n(1) = 1, prefactor 1)."""
    P = 8
    k = O.Kernel(P)
    M = k.p2m([0, 0, 0], [[0.0, 0.0, 0.5]], [1.0])
    for n in range(P):
        assert abs(M[n * (n + 1) // 2] - 0.5**n) < 1e-15
        for m in range(1, n + 1):
            assert abs(M[n * (n + 1) // 2 + m]) < 1e-15


def test_m2m_is_exact_translation():
    """P2M at the parent equals M2M of the child's P2M (no truncation upward)."""
    rng = np.random.default_rng(7)
    k = O.Kernel(9)
    Xc, Xp = np.array([0.25, 0.25, -0.25]), np.zeros(3)
    xb = _cluster(rng, Xc, 0.2, 30)
    q = rng.uniform(-1, 1, 30)
    assert _rel(k.m2m(Xp, Xc, k.p2m(Xc, xb, q)), k.p2m(Xp, xb, q)) < 1e-13


def test_l2l_is_exact_translation():
    """A local expansion shifted by L2L evaluates to the same field (L2P)."""
    rng = np.random.default_rng(8)
    k = O.Kernel(9)
    M = k.p2m([5, 5, 5], _cluster(rng, np.array([5.0, 5, 5]), 0.3, 20), rng.uniform(-1, 1, 20))
    Lp = k.m2l([0, 0, 0], [5, 5, 5], M)
    Xc = np.array([0.2, -0.1, 0.15])
    xt = _cluster(rng, Xc, 0.1, 10)
    assert _rel(k.l2p(Xc, k.l2l(Xc, [0, 0, 0], Lp), xt), k.l2p([0, 0, 0], Lp, xt)) < 1e-12


def test_m2l_converges_geometrically():
    rng = np.random.default_rng(9)
    src = _cluster(rng, np.array([3.0, 0, 0]), 0.5, 10)
    q = rng.uniform(-1, 1, 10)
    tgt = _cluster(rng, np.zeros(3), 0.5, 10)
    d = O.Kernel(4).p2p(tgt, src, q)
    errs = []
    for P in (4, 8, 12, 16):
        k = O.Kernel(P)
        L = k.m2l([0, 0, 0], [3, 0, 0], k.p2m([3, 0, 0], src, q))
        errs.append(_rel(k.l2p([0, 0, 0], L, tgt)[:, 0], d[:, 0]))
    assert all(b < a / 5 for a, b in zip(errs, errs[1:])), errs


def test_fmm_single_leaf_and_small_theta_are_direct():
    x, q = O.data_cube(50, 3), O.data_source(50, 3)
    t = O.Tree(x, 64)
    assert t.ncell == 1
    k = O.Kernel(6)
    xs, qs = x[t.perm], q[t.perm]
    assert _rel(t.fmm(k, xs, qs), O.direct(k, xs, xs, qs)) < 1e-14
    x, q = O.data_plummer(1500, 4), O.data_source(1500, 4)
    t = O.Tree(x, 16)
    ci, cj, ty = t.traverse(1e-6, 5000)
    assert int((ty == 0).sum()) == 0
    xs, qs = x[t.perm], q[t.perm]
    assert _rel(t.fmm(k, xs, qs, 1e-6), O.direct(k, xs, xs, qs)) < 1e-13


def test_fmm_linear_in_charges():
    x = O.data_cube(3000, 1)
    rng = np.random.default_rng(1)
    q1, q2 = rng.uniform(-1, 1, 3000), rng.uniform(-1, 1, 3000)
    t = O.Tree(x, 64)
    k = O.Kernel(6)
    xs = x[t.perm]
    a = t.fmm(k, xs, q1[t.perm])
    b = t.fmm(k, xs, q2[t.perm])
    c = t.fmm(k, xs, (q1 + 2 * q2)[t.perm])
    assert _rel(c, a + 2 * b) < 1e-13


def test_bounds_drift():
    """BoundBox::getBounds widens the box by 1e-5 per body (include/bound_box.h:19-20)."""
    x = np.tile([[0.5, -0.5, 0.0]], (10, 1))
    lo, hi = O.bounds(x)
    assert np.allclose(lo, x[0] - 10e-5, rtol=0, atol=1e-15)
    assert np.allclose(hi, x[0] + 10e-5, rtol=0, atol=1e-15)


def test_tree_invariants():
    x = O.data_plummer(20000, 2)
    t = O.Tree(x, 32)
    assert np.array_equal(np.sort(t.perm), np.arange(len(x)))
    leaf = t.nchild == 0
    assert t.nbody[leaf].max() <= 32 and t.nbody[leaf].sum() == len(x)
    for c in np.nonzero(~leaf)[0]:
        ch = np.arange(t.ichild[c], t.ichild[c] + t.nchild[c])
        assert np.all(t.iparent[ch] == c)
        assert t.nbody[ch].sum() == t.nbody[c]
        assert np.allclose(t.R[ch], t.R[c] / 2)


# ---------------------------------------------------------------- tree and traversal by invariants
def _cell_of_bodies(t):
    """(cell, body) index pairs for every cell and every body in its range."""
    cells = np.repeat(np.arange(t.ncell), t.nbody)
    starts = np.repeat(t.ibody, t.nbody)
    offs = np.arange(len(cells)) - np.repeat(np.cumsum(t.nbody) - t.nbody, t.nbody)
    return cells, starts + offs


def check_tree(x, t, ncrit, bnds=None):
    """What fixes BuildTree::buildTree's output (include/build_tree.h:29-209,
    bound_box.h:11-25) without running it: the root box from the drifted
    bounds, octant children at +-R/2, every body inside its cell's box and in
    the octant of its child, children in increasing octant order and
    contiguous, leaves at most ncrit bodies, and a stable split (so the bodies
    of a leaf keep their input order): together they determine the
    permutation and the cells."""
    n = len(x)
    xs = x[t.perm]
    assert np.array_equal(np.sort(t.perm), np.arange(n))
    lo, hi = O.bounds(x) if bnds is None else bnds
    X0 = (lo + hi) / 2
    R0 = max(np.max(X0 - lo), np.max(hi - X0)) * 1.00001
    assert np.array_equal(t.X[0], X0) and t.R[0] == R0
    assert t.ibody[0] == 0 and t.nbody[0] == n
    leaf = t.nchild == 0
    assert np.all(t.nbody[leaf] >= 1) and t.nbody[leaf].sum() == n
    assert np.all(t.nbody[leaf] <= ncrit)
    assert np.all(t.nbody[~leaf] > ncrit)
    # every body inside the box of every cell that holds it
    c, b = _cell_of_bodies(t)
    assert np.all(np.abs(xs[b] - t.X[c]) <= t.R[c][:, None])
    # children: contiguous block, parent link, half radius, centre at +-R/2, octants increasing
    inner = np.nonzero(~leaf)[0]
    for p in inner:
        ch = np.arange(t.ichild[p], t.ichild[p] + t.nchild[p])
        assert np.all(t.iparent[ch] == p) and np.all(t.R[ch] == t.R[p] / 2)
        d = t.X[ch] - t.X[p]
        assert np.allclose(np.abs(d), t.R[p] / 2, rtol=1e-14, atol=0)
        octs = ((d > 0) * np.array([1, 2, 4])).sum(1)
        assert np.all(np.diff(octs) > 0)
        assert t.ibody[ch[0]] == t.ibody[p] and np.all(t.ibody[ch[1:]] == t.ibody[ch[:-1]] + t.nbody[ch[:-1]])
        assert t.nbody[ch].sum() == t.nbody[p]
        # each body of a child sits in that child's octant of the parent centre
        for k, o in zip(ch, octs):
            xb = xs[t.ibody[k]:t.ibody[k] + t.nbody[k]]
            got = ((xb > t.X[p]) * np.array([1, 2, 4])).sum(1)
            assert np.all(got == o)
    # stable split: within a leaf the input order is kept
    for k in np.nonzero(leaf)[0]:
        assert np.all(np.diff(t.perm[t.ibody[k]:t.ibody[k] + t.nbody[k]]) > 0)
    # cells numbered block-depth-first (nodes2cells): the children blocks of a
    # cell's subtree follow its own children block in child order
    nxt = [1]

    def walk(cell):
        if t.nchild[cell] == 0:
            return
        assert t.ichild[cell] == nxt[0]
        nxt[0] += t.nchild[cell]
        for k in range(t.ichild[cell], t.ichild[cell] + t.nchild[cell]):
            walk(k)
    import sys
    sys.setrecursionlimit(10000)
    walk(0)
    assert nxt[0] == t.ncell


def _mac(t, a, b, theta, xper, tj=None):
    """traversal.h:117-119: the pair is far enough for M2L."""
    tj = t if tj is None else tj
    d = t.X[a] - tj.X[b] - xper
    return (d * d).sum(1) * theta * theta > (t.R[a] + tj.R[b]) ** 2 * (1 - 1e-3)


def check_lists(t, ci, cj, ty, theta, xper=np.zeros(3), tj=None):
    """What fixes the dual traversal's lists (include/traversal.h:90-177)
    without running it: every M2L pair passes the acceptance test and every
    P2P pair is two leaves that fail it; every ordered (target, source) body
    pair is covered exactly once, by a P2P pair of their leaves or an M2L pair
    of an ancestor-or-self of each; and no pair is finer than needed: one of
    the pairs it can have been split from fails the test.  tj: a separate
    source tree (traverse(icells, jcells)); ci indexes t, cj indexes tj."""
    tj = t if tj is None else tj
    leaf, leaf_j = t.nchild == 0, tj.nchild == 0
    m2l, p2p = ty == 0, ty == 1
    assert np.all(m2l | p2p)
    assert np.all(_mac(t, ci[m2l], cj[m2l], theta, xper, tj))
    assert np.all(leaf[ci[p2p]] & leaf_j[cj[p2p]])
    assert not np.any(_mac(t, ci[p2p], cj[p2p], theta, xper, tj))
    # exact cover: per target leaf, the source body ranges tile [0, n)
    nc = t.ncell
    order = np.argsort(ci, kind="stable")
    ptr = np.zeros(nc + 1, np.int64)
    ptr[1:] = np.cumsum(np.bincount(ci, minlength=nc))
    src = cj[order]
    for k in np.nonzero(leaf)[0]:
        chain, a = [k], k
        while a != 0:
            a = t.iparent[a]
            chain.append(a)
        rows = [src[ptr[a]:ptr[a + 1]] for a in chain]
        # P2P sources only on the leaf itself; ancestors hold M2L sources only
        s = np.concatenate(rows)
        st, ln = tj.ibody[s], tj.nbody[s]
        o = np.argsort(st, kind="stable")
        st, ln = st[o], ln[o]
        assert st[0] == 0 and np.all(st[1:] == st[:-1] + ln[:-1]) and st[-1] + ln[-1] == tj.n, k
        for a in chain[1:]:
            assert np.all(ty[order[ptr[a]:ptr[a + 1]]] == 0)
    # not finer than needed
    root = (ci == 0) & (cj == 0)
    a, b = ci[~root], cj[~root]
    pa, pb = t.iparent[a], tj.iparent[b]
    has_pa, has_pb = a != 0, b != 0
    fail_a = has_pa & ~_mac(t, np.where(has_pa, pa, 0), b, theta, xper, tj)
    fail_b = has_pb & ~_mac(t, a, np.where(has_pb, pb, 0), theta, xper, tj)
    fail_ab = has_pa & has_pb & ~_mac(t, np.where(has_pa, pa, 0), np.where(has_pb, pb, 0), theta, xper, tj)
    assert np.all(fail_a | fail_b | fail_ab)


@pytest.mark.parametrize("dist,n,ncrit", [("cube", 20000, 64), ("plummer", 15000, 16), ("cube", 3000, 4),
                                          ("plummer", 40000, 64)])
def test_tree_fixed_by_invariants(dist, n, ncrit):
    x = O.data_cube(n, 7) if dist == "cube" else O.data_plummer(n, 7)
    check_tree(x, O.Tree(x, ncrit), ncrit)


@pytest.mark.parametrize("dist,n,ncrit,theta,nspawn,shift", [
    ("cube", 20000, 64, 0.4, 5000, 0), ("plummer", 15000, 16, 0.3, 5000, 0), ("cube", 8000, 32, 0.6, 100, 0),
    ("plummer", 10000, 8, 0.5, 1, 0), ("cube", 12000, 64, 0.35, 5000, 0), ("cube", 6000, 32, 0.4, 5000, 1),
    ("plummer", 30000, 64, 0.4, 5000, 0)])
def test_lists_fixed_by_invariants(dist, n, ncrit, theta, nspawn, shift):
    x = O.data_cube(n, 3) if dist == "cube" else O.data_plummer(n, 3)
    t = O.Tree(x, ncrit)
    xper = np.array([2 * math.pi, 0.0, -2 * math.pi]) * shift
    ci, cj, ty = t.traverse(theta, nspawn, xper)
    check_lists(t, ci, cj, ty, theta, xper)


@pytest.mark.parametrize("sign", [1, -1])
def test_acceptance_test_known_answer(sign):
    """Two bodies at opposite corners, ncrit = 1: the root splits into two
    leaves of radius R0/2 whose centres differ by R0 along every axis, so
    |dX|^2 = 3 R0^2 and (Ri + Rj)^2 = R0^2.  The acceptance test
    |dX|^2 theta^2 > (Ri + Rj)^2 (1 - 1e-3) (traversal.h:117-119) flips at
    theta^2 = (1 - 1e-3) / 3: just above it the two cross pairs are M2L, just
    below they are P2P; between that and 1/3 (where a test without the 1e-3
    margin would flip) they are still M2L."""
    x = np.array([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]])
    t = O.Tree(x, 1)
    assert t.ncell == 3 and np.array_equal(t.X[2] - t.X[1], np.full(3, t.R[0]))
    th = math.sqrt((1 - 1e-3) / 3) * (1 + sign * 1e-9)
    ci, cj, ty = t.traverse(th, 5000)
    assert len(ty) == 4 and sorted(zip(ci, cj)) == [(1, 1), (1, 2), (2, 1), (2, 2)]
    cross = ci != cj
    assert np.all(ty[~cross] == 1)
    assert np.all(ty[cross] == (0 if sign > 0 else 1))
    _, _, ty = t.traverse(math.sqrt(1 / 3) * (1 - 1e-5), 5000)
    assert int((ty == 0).sum()) == 2


def _ine_j(n, ncrit, dist="cube"):
    """examples/fmm.cxx -j: both trees in the bounds of all bodies (:59-62)."""
    xi, qi, xj, qj = O.ine_j_bodies(n, dist=dist)
    b = O.bounds(xj, O.bounds(xi))
    return xi, qi, xj, qj, O.Tree(xi, ncrit, b), O.Tree(xj, ncrit, b)


@pytest.mark.parametrize("dist,n,ncrit,theta,nspawn", [("cube", 8000, 64, 0.4, 5000),
                                                       ("plummer", 6000, 16, 0.5, 100),
                                                       ("cube", 3000, 8, 0.3, 1)])
def test_pair_lists_fixed_by_invariants(dist, n, ncrit, theta, nspawn):
    """traverse(icells, jcells): target bodies against source bodies only,
    every pair once, by the same acceptance and split rules."""
    xi, qi, xj, qj, ti, tj = _ine_j(n, ncrit, dist)
    if n < 4000:
        b = O.bounds(xj, O.bounds(xi))
        check_tree(xi, ti, ncrit, b)
        check_tree(xj, tj, ncrit, b)
    ci, cj, ty = ti.traverse_pair(tj, theta, nspawn)
    assert len(ci) > 0
    check_lists(ti, ci, cj, ty, theta, tj=tj)


def test_pair_fmm_converges_to_direct():
    """The pair solve against the direct sum of the sources at the targets:
    theta -> 0 makes it P2P only (exact to rounding), and at theta = 0.4 the
    error falls geometrically with P."""
    xi, qi, xj, qj, ti, tj = _ine_j(3000, 32)
    xs, xjs, qjs = xi[ti.perm], xj[tj.perm], qj[tj.perm]
    k = O.Kernel(4)
    d = O.direct(k, xs, xjs, qjs)
    assert _rel(ti.fmm_pair(k, xs, tj, xjs, qjs, 1e-6), d) < 1e-13
    errs = [_rel(ti.fmm_pair(O.Kernel(P), xs, tj, xjs, qjs, 0.4)[:, 0], d[:, 0]) for P in (4, 8, 12)]
    assert errs[0] < 1e-2 and errs[1] < errs[0] / 10 and errs[2] < errs[1] / 10, errs


def test_invariant_checks_catch_planted_mistakes():
    """The checks above are not vacuous: lists with a wrong acceptance factor,
    a dropped pair or a duplicated pair, and a tree with two bodies swapped,
    each fail them."""
    x = O.data_cube(6000, 4)
    t = O.Tree(x, 32)
    ci, cj, ty = t.traverse(0.4, 5000)
    check_lists(t, ci, cj, ty, 0.4)
    bad = [
        t.traverse(0.4 * 1.1, 5000),                           # looser acceptance
        (np.delete(ci, 100), np.delete(cj, 100), np.delete(ty, 100)),   # dropped pair
        (np.append(ci, ci[7]), np.append(cj, cj[7]), np.append(ty, ty[7])),  # duplicated pair
    ]
    for lists in bad:
        with pytest.raises(AssertionError):
            check_lists(t, *lists, 0.4)
    check_tree(x, t, 32)
    leaf = np.nonzero(t.nchild == 0)[0][3]
    i = t.ibody[leaf]
    t.perm[[i, i + 1]] = t.perm[[i + 1, i]]
    with pytest.raises(AssertionError):
        check_tree(x, t, 32)


# ---------------------------------------------------------------- periodic images
def _image_sum(k, xs, qs, shifts):
    """Brute force: every target against every source shifted by each of `shifts`."""
    out = np.zeros((len(xs), 4))
    for s in shifts:
        out += O.direct(k, xs, xs + np.asarray(s), qs)
    return out


def test_periodic_one_image_level_is_the_27_box_sum():
    """images = 1: the dual traversal over the 27 shifts (include/traversal.h:335-347)
    and no supercells: the FMM of the 3x3x3 copies."""
    x, q = O.data_cube(300, 4), O.data_source(300, 4)
    t = O.Tree(x, 32)
    xs, qs = x[t.perm], q[t.perm]
    k = O.Kernel(10)
    cyc = 2 * np.pi
    trg = t.fmm_periodic(k, xs, qs, 0.4, 5000, 1, cyc)
    shifts = [(ix * cyc, iy * cyc, iz * cyc) for ix in (-1, 0, 1) for iy in (-1, 0, 1) for iz in (-1, 0, 1)]
    d = _image_sum(k, xs, qs, shifts)
    assert _rel(trg[:, 0], d[:, 0]) < 1e-5 and _rel(trg[:, 1:], d[:, 1:]) < 1e-4


def test_periodic_supercell_level_adds_its_boxes():
    """images = 2: traversePeriodic (include/traversal.h:233-287) adds the root's
    copies at (i*nb + c) * cycle, i != 0, c in [-prange, prange]^3 (prange =
    int(.5/theta + 1), nb = 2 prange + 1) by M2L; checked against brute force
    over exactly the near 27 and those boxes."""
    x, q = O.data_cube(120, 5), O.data_source(120, 5)
    t = O.Tree(x, 32)
    xs, qs = x[t.perm], q[t.perm]
    k = O.Kernel(10)
    cyc, theta = 2 * np.pi, 0.4
    pr = int(0.5 / theta + 1)  # 2: the shell at distance 2 belongs to neither part
    nb = 2 * pr + 1
    rng = range(-pr, pr + 1)
    shifts = [(ix * cyc, iy * cyc, iz * cyc) for ix in (-1, 0, 1) for iy in (-1, 0, 1) for iz in (-1, 0, 1)]
    shifts += [((ix * nb + cx) * cyc, (iy * nb + cy) * cyc, (iz * nb + cz) * cyc)
               for ix in rng for iy in rng for iz in rng if (ix, iy, iz) != (0, 0, 0)
               for cx in rng for cy in rng for cz in rng]
    trg = t.fmm_periodic(k, xs, qs, theta, 5000, 2, cyc)
    d = _image_sum(k, xs, qs, shifts)
    assert _rel(trg[:, 0], d[:, 0]) < 1e-5 and _rel(trg[:, 1:], d[:, 1:]) < 1e-4
    # the supercell part alone, against its boxes
    near = t.fmm_periodic(k, xs, qs, theta, 5000, 1, cyc)
    far = _image_sum(k, xs, qs, shifts[27:])
    assert _rel((trg - near)[:, 0], far[:, 0]) < 3e-5 and _rel((trg - near)[:, 1:], far[:, 1:]) < 3e-5


# ---------------------------------------------------------------- Ewald and dipole correction
MADELUNG_NACL = 1.747564594633182  # rock salt, per ion, nearest-neighbour distance 1


def _rock_salt(m):
    """m^3 ions of alternating charge on the unit lattice, box m (m even), D = 0"""
    g = np.arange(m)
    i, j, k = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([i, j, k], -1).reshape(-1, 3).astype(float) - (m - 1) / 2
    q = np.where((i + j + k).reshape(-1) % 2 == 0, 1.0, -1.0)
    return x, q


def test_ewald_rock_salt_madelung_constant():
    """Ewald potential of the rock-salt lattice: q_i phi_i = -M for every ion."""
    x, q = _rock_salt(4)
    L = 4.0
    assert np.allclose((x * q[:, None]).sum(0), 0)
    trg = O.ewald(x, q, ksize=24, alpha=3.0, sigma=0.25 / np.pi, cutoff=L / 2, cycle=L)
    assert np.max(np.abs(q * trg[:, 0] + MADELUNG_NACL)) < 1e-10
    assert np.max(np.abs(trg[:, 1:])) < 1e-10  # every ion is a centre of symmetry


def _neutral(n, L, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)) * L - L / 2
    q = rng.random(n) - 0.5
    return x, q - q.mean()


def test_ewald_split_parameter_does_not_matter():
    """alpha only moves work between the real and wave sums (both converged)."""
    L = 2 * np.pi
    x, q = _neutral(150, L, 1)
    a = O.ewald(x, q, 17, 10 / L, 0.25 / np.pi, L / 2, L)
    b = O.ewald(x, q, 21, 12 / L, 0.25 / np.pi, L / 2, L)
    assert _rel(a[:, 0], b[:, 0]) < 1e-9 and _rel(a[:, 1:], b[:, 1:]) < 1e-9


def test_ewald_gradient_is_energy_gradient():
    """dE/dx_i = q_i grad phi_i for E = 1/2 sum q phi (central differences)."""
    L = 2 * np.pi
    x, q = _neutral(40, L, 2)
    args = (14, 8 / L, 0.25 / np.pi, L / 2, L)
    g = O.ewald(x, q, *args)[:, 1:]
    h = 1e-5
    for i in (0, 17, 33):
        for d in range(3):
            e = []
            for s in (h, -h):
                xp = x.copy()
                xp[i, d] += s
                e.append(0.5 * q @ O.ewald(xp, q, *args)[:, 0])
            fd = (e[0] - e[1]) / (2 * h)
            assert abs(fd - q[i] * g[i, d]) < 1e-6 * max(1.0, abs(q[i] * g[i, d]))


def test_dipole_correction_turns_lattice_sum_into_ewald():
    """A cube of (2K+1)^3 images (vacuum boundary) minus the dipole term of
    include/up_down_pass.h:89-99 approaches the tin-foil Ewald sum as O(1/K^2);
    without the correction it does not."""
    L = 2 * np.pi
    x, q = _neutral(50, L, 3)
    ew = O.ewald(x, q, 17, 10 / L, 0.25 / np.pi, L / 2, L)
    k = O.Kernel(4)
    K = 8
    s = np.zeros((len(q), 4))
    rng = range(-K, K + 1)
    for a in rng:
        for b in rng:
            for c in rng:
                s += O.direct(k, x, x + np.array([a, b, c]) * L, q)
    raw = s.copy()
    O.dipole_correction(x, q, L, s)
    energy = lambda t: q @ t[:, 0]
    assert abs(energy(s) - energy(ew)) < 2e-3 * abs(energy(ew))
    assert _rel(s[:, 1:], ew[:, 1:]) < 3e-4
    assert _rel(raw[:, 1:], ew[:, 1:]) > 1e-2


def test_wrap_into_box():
    L = np.array([2.0, 3.0, 4.0])
    x = np.array([[0.5, -1.2, 2.5], [-1.5, 1.6, -2.1], [1.0, 1.5, 2.0]])
    w, iw = O.wrap(x, L)
    assert np.allclose(w, [[0.5, -1.2, -1.5], [0.5, -1.4, 1.9], [1.0, 1.5, 2.0]])
    assert list(iw) == [4, 7, 0]
