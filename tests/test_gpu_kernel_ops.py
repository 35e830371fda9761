"""The per-cell Kernel operators (kernels/laplace.h:150-388) of the drop-in,
through the C ABI (fmm_kernel_*) and the C++ layer, and the separate source
tree path (Traversal::traverse(icells, jcells), examples/fmm.cxx -j).

Bars: operators against the oracle's (themselves equal to the reference's
compiled operators) at 1e-12 relative; the examples/kernel.cxx chain against
the fixture written from the reference's operators at 1e-12 (the direct P2P
at 1e-14); the pair solve against the oracle's two-tree FMM at 1e-10.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "exafmm-beta_amd")
G = os.path.join(ROOT, "tests", "golden")
OP_TOL = 1e-12


def _F():
    import exafmm_amd
    return exafmm_amd


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("P", [3, 8, 12, 20])
def test_operators_match_oracle(P):
    F = _F()
    rng = np.random.default_rng(P)
    k = O.Kernel(P)
    f = F.FMM(P=P)
    Xc = rng.uniform(-1, 1, 3)
    xb = Xc + rng.uniform(-0.5, 0.5, (37, 3))
    q = rng.uniform(-1, 1, 37)
    # P2M, accumulating into a non-zero M
    M0 = rng.normal(size=k.nterm) + 1j * rng.normal(size=k.nterm)
    M = f.kernel_p2m(Xc, xb, q, M0.copy())
    assert _rel(M - M0, k.p2m(Xc, xb, q)) < OP_TOL
    # M2M of three children into a parent, in a tree's geometry (children at
    # half the parent's radius per axis, bodies within a child's radius): the
    # oracle's, and the exact identity M2M(P2M about the children) = P2M about
    # the parent (M2M of degrees below P is exact)
    Xp = rng.uniform(-1, 1, 3)
    kids = [Xp + 0.5 * np.array(o) for o in ([1, -1, 1], [-1, -1, -1], [-1, 1, 1])]
    xbs = [c + rng.uniform(-0.5, 0.5, (37, 3)) for c in kids]
    qbs = [rng.uniform(-1, 1, 37) for c in kids]
    Ms = [k.p2m(c, xk, qk) for c, xk, qk in zip(kids, xbs, qbs)]
    ref = np.zeros(k.nterm, complex)
    for c, m in zip(kids, Ms):
        k.m2m(Xp, c, m, ref)
    got = f.kernel_m2m(Xp, np.array(kids), np.array(Ms))
    assert _rel(got, ref) < OP_TOL
    assert _rel(got, k.p2m(Xp, np.concatenate(xbs), np.concatenate(qbs))) < 1e-13
    # M2L, batched, with and without a periodic shift
    Mp = ref
    Xi = np.array([Xp + [4.0, 1.0, -3.0], Xp + [-5.0, 2.0, 1.0], Xp + [0.5, 6.0, 0.0]])
    Xj = np.array([Xp, Xp, Xp])
    Ls = f.kernel_m2l(Xi, Xj, [Mp, Mp, Mp])
    for r in range(3):
        assert _rel(Ls[r], k.m2l(Xi[r], Xj[r], Mp)) < OP_TOL
    xper = np.array([2.0, -1.0, 0.5])
    Lsh = f.kernel_m2l(Xi[:1], Xj[:1] - xper, [Mp], xper=xper)
    assert _rel(Lsh[0], k.m2l(Xi[0], Xj[0] - xper, Mp, xper=xper)) < OP_TOL
    assert _rel(Lsh[0], Ls[0]) < 1e-12
    # L2L, L2P
    Xl = Xi[0] + np.array([-0.25, 0.5, 0.25])
    Ll = f.kernel_l2l(Xl, Xi[0], Ls[0])
    assert _rel(Ll, k.l2l(Xl, Xi[0], Ls[0])) < OP_TOL
    xt = Xl + rng.uniform(-0.25, 0.25, (21, 3))
    assert _rel(f.kernel_l2p(Xl, Ll, xt), k.l2p(Xl, Ll, xt)) < OP_TOL
    # P2P with a coincident pair (skipped) and a shift
    xs = np.concatenate([xb, xt[:2]])
    qs = np.concatenate([q, [0.3, -0.2]])
    got = f.kernel_p2p(xt, xs, qs)
    assert _rel(got, k.p2p(xt, xs, qs)) < OP_TOL and np.all(np.isfinite(got))
    assert _rel(f.kernel_p2p(xt, xs, qs, xper=xper), k.p2p(xt, xs, qs, xper=xper)) < OP_TOL
    f.close()


def test_kernel_example_matches_fixture():
    """examples/kernel.cxx's chain (:35-86) run through the C++ layer's Kernel
    methods: the FMM value at the target and the direct P2P value equal the
    fixture written from the reference's operators."""
    fx = np.load(os.path.join(G, "chain.npz"))
    exe = os.path.join(PKG, "lib", "kernel")
    for P in (4, 10, 20):
        out = subprocess.run([exe, "-P", str(P)], capture_output=True, text=True, timeout=60)
        assert out.returncode == 0, out.stderr
        rows = dict((m[0], np.array(m[1:], float)) for m in
                    (l.split() for l in out.stdout.splitlines() if l.startswith(("fmm ", "direct "))))
        assert _rel(rows["fmm"], fx["fmm_p%d" % P][0]) < 1e-12, P
        assert _rel(rows["direct"], fx["direct_p%d" % P][0]) < 1e-14, P
        first = out.stdout.splitlines()[0].split()
        assert int(first[0]) == P
        # Verify::getDifScalar / getNrmScalar(bodies): normalised by the FMM value
        fm, di = fx["fmm_p%d" % P][0][0], fx["direct_p%d" % P][0][0]
        err = abs(fm - di) / abs(fm)
        assert abs(float(first[1]) - err) <= 1e-5 * err


# the last column bounds the error against the direct sum: the truncation
# error of the two-tree FMM with zero-mean charges (the oracle's own:
# cube P = 8, theta = 0.4: 4e-4; Plummer P = 6, theta = 0.5: 1.7e-2 / 2.3e-2,
# 1.8e-3 / 2.0e-3 at P = 10), with a margin of about 2x
@pytest.mark.parametrize("dist,n,ncrit,theta,P,bound", [("cube", 20000, 64, 0.4, 8, 1e-3),
                                                        ("plummer", 8000, 16, 0.5, 6, 5e-2)])
def test_pair_solve_matches_oracle(dist, n, ncrit, theta, P, bound):
    """Targets against a separate source tree (fmm_plan_pair): every target
    body against the oracle's two-tree FMM at 1e-10; the source bodies get
    nothing."""
    F = _F()
    xi, qi, xj, qj = O.ine_j_bodies(n, dist=dist)
    lo, hi = O.bounds(xj, O.bounds(xi))
    pi = F.Plan.tree_in_bounds(xi, ncrit, lo, hi)
    pj = F.Plan.tree_in_bounds(xj, ncrit, lo, hi)
    pp = F.Plan.pair(pi, pj, theta, 5000)
    f = F.FMM(P=P, ncrit=ncrit, theta=theta)
    f.set_bodies_plan(pp, np.concatenate([qi, qj]))
    f.evaluate()
    trg = f.targets(input_order=False)
    f.close()
    ti, tj = O.Tree(xi, ncrit, (lo, hi)), O.Tree(xj, ncrit, (lo, hi))
    ref = ti.fmm_pair(O.Kernel(P), xi[ti.perm], tj, xj[tj.perm], qj[tj.perm], theta, 5000)
    assert _rel(trg[:n, 0], ref[:, 0]) < 1e-10 and _rel(trg[:n, 1:], ref[:, 1:]) < 1e-10
    assert np.all(trg[n:] == 0)
    d = O.direct(O.Kernel(P), xi[ti.perm][::97], xj, qj)
    e = O.rel_l2(d, trg[:n][::97])
    assert e[0] < bound and e[1] < bound, e


def test_example_driver_separate_sources():
    """examples/fmm.cxx -j: targets of seed 0 in x in [0, pi], sources of seed 1
    in x in [-pi, 0], both trees in the joint bounds; the printed errors equal
    the oracle's two-tree FMM against the direct sum of the sources at the 100
    sampled targets."""
    n, P = 20000, 10
    exe = os.path.join(PKG, "lib", "fmm")
    out = subprocess.run([exe, "-n", str(n), "-j", "-v"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = {k: float(v) for k, v in re.findall(r"Rel\. L2 Error \((pot|acc)\)\s*:\s*(\S+)", out.stdout)}
    xi, qi, xj, qj = O.ine_j_bodies(n)
    lo, hi = O.bounds(xj, O.bounds(xi))
    ti, tj = O.Tree(xi, 64, (lo, hi)), O.Tree(xj, 64, (lo, hi))
    k = O.Kernel(P)
    fmm = ti.fmm_pair(k, xi[ti.perm], tj, xj[tj.perm], qj[tj.perm], 0.4, 5000)
    idx = O.sample_indices(n, 100)
    d = O.direct(k, xi[ti.perm][idx], xj[tj.perm], qj[tj.perm])
    pot, acc = O.rel_l2(d, fmm[idx])
    print("-j: printed %s, oracle %.7e %.7e" % (got, pot, acc))
    assert abs(got["pot"] - pot) <= 1e-6 * pot and abs(got["acc"] - acc) <= 1e-6 * acc
    assert pot < 1e-4 and acc < 1e-4


def test_direct_keeps_the_solve_uploaded(tmp_path):
    """Traversal::direct runs in a context of its own: a downward pass after
    it still finds the tree of the upward pass (tests/cpp/direct_then_downward.cpp:
    upward, traverse, direct, downward), and the FMM agrees with that direct sum."""
    F = _F()
    exe = str(tmp_path / "direct_then_downward")
    libdir = os.path.dirname(F.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "direct_then_downward.cpp"), "-o", exe,
                    "-L", libdir, "-lexafmm_amd", "-Wl,-rpath," + libdir], check=True)
    out = subprocess.run([exe, "-n", "8192"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    m = re.search(r"pot (\S+) acc (\S+)", out.stdout)
    assert m and float(m.group(1)) < 1e-4 and float(m.group(2)) < 1e-4, out.stdout
