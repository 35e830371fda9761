"""CPU checks of the product's host side (through the C ABI of
libexafmm_amd.so): exported symbols, and the tree / interaction lists built by
fmm_plan_build against the oracle (identical trees, identical pair lists in
the reference's visiting order)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _F():
    import exafmm_amd
    return exafmm_amd


def test_library_exports_every_declared_symbol():
    F = _F()
    hdr = open(os.path.join(ROOT, "include", "exafmm_amd.h")).read()
    names = set(re.findall(r"\b(fmm_[a-z_0-9]+)\s*\(", hdr))
    assert names, "no declarations found"
    lib = ctypes.CDLL(F.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(F.SYMBOLS) | (names - set(F.SYMBOLS))
    assert set(F.SYMBOLS) <= names


def _oracle_csr(t, ci, cj, ty, typ):
    sel = ty == typ
    a, b = ci[sel], cj[sel]
    order = np.argsort(a, kind="stable")
    ptr = np.zeros(t.ncell + 1, np.int64)
    ptr[1:] = np.cumsum(np.bincount(a, minlength=t.ncell))
    return ptr, b[order]


@pytest.mark.parametrize("dist,n,ncrit,theta,nspawn", [
    ("cube", 32768, 64, 0.4, 5000), ("plummer", 20000, 64, 0.4, 5000),
    ("plummer", 20000, 8, 0.25, 100), ("cube", 5000, 16, 0.6, 1000), ("cube", 1, 64, 0.4, 5000),
    ("cube", 64, 64, 0.4, 5000),
    # above 50000 bodies the lists are built by several threads
    ("plummer", 120000, 64, 0.4, 5000), ("cube", 150000, 32, 0.5, 1000)])
def test_plan_matches_oracle(dist, n, ncrit, theta, nspawn):
    F = _F()
    x = O.data_cube(n, 2) if dist == "cube" else O.data_plummer(n, 2)
    pl = F.Plan(x, ncrit=ncrit, theta=theta, nspawn=nspawn)
    t = O.Tree(x, ncrit)
    tr = pl.tree()
    assert np.array_equal(tr["perm"], t.perm)
    for k in ("iparent", "ichild", "nchild", "ibody", "nbody"):
        assert np.array_equal(tr[k], getattr(t, k)), k
    assert np.array_equal(tr["X"], t.X) and np.array_equal(tr["R"], t.R)
    assert tr["depth"] == t.depth
    ci, cj, ty = t.traverse(theta, nspawn)
    mp, ms, pp, ps = pl.lists()
    optr, osrc = _oracle_csr(t, ci, cj, ty, 0)
    assert np.array_equal(mp, optr) and np.array_equal(ms, osrc)
    optr, osrc = _oracle_csr(t, ci, cj, ty, 1)
    assert np.array_equal(pp, optr) and np.array_equal(ps, osrc)
    s = pl.stats()
    assert s["m2l_pairs"] == int((ty == 0).sum()) and s["p2p_pairs"] == int((ty == 1).sum())


@pytest.mark.parametrize("case", ["cube", "first_far", "tiny", "huge", "plummer"])
def test_bounds_chained_scan_equals_sequential(case):
    """BoundBox::getBounds (include/bound_box.h:16-23) widens the box by 1e-5
    per body, rounding each step.  fmm_plan_build scans runs of bodies on
    host threads (n >= 65536) and chains the runs with an exact k-fold shift;
    the box must equal the oracle's one sequential scan bit for bit, also
    where the drift of the first body decides it (crossing zero, small
    binades, coordinates whose spacing exceeds the step)."""
    F = _F()
    rng = np.random.default_rng(11)
    n = 70001 if case != "plummer" else 131072
    if case == "plummer":
        x, _ = F.make_bodies(n, "plummer1", 2)
    else:
        x = rng.uniform(-1, 1, (n, 3))
    if case == "first_far":
        x = 1.0 + np.abs(x)
        x[0] = (-0.5, 0.7, 3.0517578125e-05)
    elif case == "tiny":
        x *= 1e-3
    elif case == "huge":
        x = x * 1e11 + 3e11
    lo, hi = O.bounds(np.ascontiguousarray(x))
    s = F.Plan(x, ncrit=64).stats()
    assert np.array_equal(np.array(s["xmin"]), np.asarray(lo)), (s["xmin"], lo)
    assert np.array_equal(np.array(s["xmax"]), np.asarray(hi)), (s["xmax"], hi)


def test_plan_empty_and_duplicates():
    F = _F()
    pl = F.Plan(np.zeros((0, 3)))
    assert pl.stats()["ncell"] == 0
    x = O.data_cube(500, 3)
    x[100:140] = x[0]  # 41 coincident bodies (<= ncrit): same tree as the reference
    pl = F.Plan(x, ncrit=64)
    t = O.Tree(x, 64)
    assert pl.stats()["ncell"] == t.ncell
    assert np.array_equal(pl.tree()["perm"], t.perm)
    # more than ncrit coincident bodies never separate: the reference recursion
    # does not terminate there; the product stops at level 30 with one leaf
    x[100:300] = x[0]
    tr = F.Plan(x, ncrit=64).tree()
    assert tr["depth"] == 30
    assert tr["nbody"][tr["nchild"] == 0].max() == 201


def test_make_bodies_matches_generators():
    F = _F()
    x, q = F.make_bodies(4000, "cube", 3)
    assert np.array_equal(x, O.data_cube(4000, 3)) and np.array_equal(q, O.data_source(4000, 3))
    x, q = F.make_bodies(4000, "plummer", 3)
    assert np.array_equal(x, O.data_plummer(4000, 3)) and np.array_equal(q, O.data_source(4000, 3))
    if O.ref_available():
        for d in ("cube", "sphere"):
            a, b = F.make_bodies(4096, d, 1)
            c, e = O.ref_init_bodies(4096, d, 1)
            assert np.array_equal(a, c) and np.array_equal(b, e), d
    if O.ref_ieee_available():
        # the Plummer and octant draws round differently under the reference's
        # -ffast-math (last bits only); without it they are bit-equal
        for d in ("cube", "sphere", "octant", "plummer"):
            a, b = F.make_bodies(4096, d, 1)
            c, e = O.ref_init_bodies(4096, d, 1, ieee=True)
            assert np.array_equal(a, c) and np.array_equal(b, e), d
    x, q = F.make_bodies(10000, "plummer1", 0)
    assert np.all(np.linalg.norm(x, axis=1) < 10) and np.all(q == 1e-4)
    x, q = F.make_bodies(10000, "cube1", 0)
    assert np.all(np.abs(x) <= 1) and abs(q.mean()) < 1e-15 and np.abs(q).max() <= 1.001
    with pytest.raises(F.FMMError):
        F.make_bodies(10, "nonsense", 0)


def test_create_reports_errors():
    F = _F()
    with pytest.raises(F.FMMError):
        F.FMM(P=1)            # unsupported order, rejected before any device call
    with pytest.raises(F.FMMError):
        F.FMM(P=10, theta=0)  # invalid parameter
    with pytest.raises(F.FMMError):
        F.FMM(P=14, precision=F.FMM_MIXED)  # FP32 M2L is built for P <= 12
    with pytest.raises(F.FMMError):
        F.FMM(P=8, precision=F.FMM_MIXED, m2l_mode=F.M2L_TARGET)  # mixed needs grouped M2L


def test_plan_merge_is_a_forest():
    F = _F()
    xs = [O.data_cube(3000, 1), O.data_plummer(5000, 2), O.data_cube(40, 3)]
    plans = [F.Plan(x, ncrit=32) for x in xs]
    for p in plans:
        p.build_groups(F.FMM_MIXED)
    m = F.Plan.merge(plans)
    tr, mp, ms, pp, ps = m.tree(), *m.lists()
    bo = co = 0
    for x, p in zip(xs, plans):
        t = p.tree()
        nc, n = len(t["R"]), len(x)
        assert np.array_equal(tr["perm"][bo:bo + n], t["perm"] + bo)
        for k in ("iparent", "ichild", "ibody"):
            off = bo if k == "ibody" else co
            assert np.array_equal(tr[k][co:co + nc], t[k] + off), k
        assert np.array_equal(tr["nbody"][co:co + nc], t["nbody"])
        assert np.array_equal(tr["X"][co:co + nc], t["X"]) and np.array_equal(tr["R"][co:co + nc], t["R"])
        a, b, c, d = p.lists()
        assert np.array_equal(np.diff(mp[co:co + nc + 1]), np.diff(a))
        assert np.array_equal(ms[mp[co]:mp[co + nc]], b + co)
        assert np.array_equal(ps[pp[co]:pp[co + nc]], d + co)
        bo += n
        co += nc
    s = m.stats()
    assert s["n"] == sum(len(x) for x in xs) and s["ncell"] == co
    assert s["depth"] == max(p.stats()["depth"] for p in plans)
    assert s["p2p_body_pairs"] == sum(p.stats()["p2p_body_pairs"] for p in plans)


def test_server_without_device_reports_error():
    """No GPU in this process: the server refuses to start with an error code
    (its parameters are checked with a context first) instead of crashing."""
    F = _F()
    import os
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible here")
    with pytest.raises(F.FMMError):
        F.Server(P=8)
    with pytest.raises(F.FMMError):
        F.Server(P=1)


@pytest.mark.parametrize("n,tt", [(20000, 32), (60000, 64)])
def test_m2l_tiles_cover_every_pair(tmp_path, n, tt):
    """tests/cpp/tiles_check.cpp against the package's host objects."""
    import subprocess
    pkg = os.path.join(ROOT, "exafmm-beta_amd")
    exe = str(tmp_path / "tiles_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(pkg, "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "tiles_check.cpp"),
                    os.path.join(pkg, "build", "fmm_host.o"), "-o", exe], check=True)
    out = subprocess.run([exe, str(n), str(tt)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


@pytest.mark.parametrize("dist,n,ncrit,theta,nspawn", [("cube", 8000, 64, 0.4, 5000),
                                                       ("plummer", 6000, 16, 0.5, 100),
                                                       ("cube", 120000, 64, 0.4, 5000)])
def test_pair_plan_matches_oracle(dist, n, ncrit, theta, nspawn):
    """fmm_plan_build_tree in shared bounds + fmm_plan_pair (traverse(icells,
    jcells), the -j path of examples/fmm.cxx): the forest holds both trees
    exactly as the oracle builds them, and the pair lists equal the oracle's
    two-tree traversal (source cells offset by the target tree's cell count)."""
    F = _F()
    xi, qi, xj, qj = O.ine_j_bodies(n, dist=dist)
    lo, hi = O.bounds(xj, O.bounds(xi))
    ti, tj = O.Tree(xi, ncrit, (lo, hi)), O.Tree(xj, ncrit, (lo, hi))
    pi = F.Plan.tree_in_bounds(xi, ncrit, lo, hi)
    pj = F.Plan.tree_in_bounds(xj, ncrit, lo, hi)
    for p, t in ((pi, ti), (pj, tj)):
        tr = p.tree()
        assert np.array_equal(tr["perm"], t.perm)
        for k in ("iparent", "ichild", "nchild", "ibody", "nbody"):
            assert np.array_equal(tr[k], getattr(t, k)), k
        assert np.array_equal(tr["X"], t.X) and np.array_equal(tr["R"], t.R)
        assert p.stats()["m2l_pairs"] == 0 and p.stats()["p2p_pairs"] == 0
    pp = F.Plan.pair(pi, pj, theta, nspawn)
    s = pp.stats()
    assert s["n"] == 2 * n and s["ncell"] == ti.ncell + tj.ncell
    ci, cj, ty = ti.traverse_pair(tj, theta, nspawn)
    mp, ms, ptr, ps = pp.lists()
    nci = ti.ncell
    for typ, (p_, s_) in ((0, (mp, ms)), (1, (ptr, ps))):
        sel = ty == typ
        a, b = ci[sel], cj[sel] + nci
        order = np.argsort(a, kind="stable")
        optr = np.zeros(nci + tj.ncell + 1, np.int64)
        optr[1:nci + 1] = np.cumsum(np.bincount(a, minlength=nci))
        optr[nci + 1:] = optr[nci]
        assert np.array_equal(p_, optr) and np.array_equal(s_, b[order])
    assert s["m2l_pairs"] == int((ty == 0).sum()) and s["p2p_pairs"] == int((ty == 1).sum())
