"""The C++ drop-in layer (include/exafmm_amd.hpp) used the way the reference's
examples/fmm.cxx uses its classes.

CPU: a host-only program (tests/cpp/tree_dump.cpp: Args, Dataset, BoundBox,
BuildTree) is compiled with g++ and its bodies and cells are compared with the
oracle's generator and tree.  GPU: the example driver (examples/fmm.cpp)
prints the same "Rel. L2 Error" lines as the reference driver, checked against
the reference's own printed values; fmm_set_tree with a caller-built tree
gives the same result as fmm_set_bodies.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "exafmm-beta_amd")


def _F():
    import exafmm_amd
    return exafmm_amd


@pytest.fixture(scope="module")
def tree_dump(tmp_path_factory):
    F = _F()
    exe = str(tmp_path_factory.mktemp("cpp") / "tree_dump")
    libdir = os.path.dirname(F.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "tree_dump.cpp"), "-o", exe,
                    "-L", libdir, "-lexafmm_amd", "-Wl,-rpath," + libdir], check=True)
    return exe


def _run_dump(exe, *args):
    out = subprocess.run([exe, *args], check=True, capture_output=True, text=True).stdout.split("\n")
    head = out[0].split()
    lo_hi = np.array(out[1].split(), float)
    nb, nc = map(int, out[2].split())
    bodies = np.array([l.split() for l in out[3:3 + nb]], float).reshape(nb, 5)
    cells = np.array([l.split()[:10] for l in out[3 + nb:3 + nb + nc]], float).reshape(nc, 10)
    return head, lo_hi, bodies, cells


@pytest.mark.parametrize("argv,n,ncrit,dist", [
    ([], 1000000, 64, "cube"),
    (["-n", "5000", "-c", "16", "-d", "plummer"], 5000, 16, "plummer"),
    (["--numBodies", "3000", "--ncrit", "8", "--distribution", "c"], 3000, 8, "cube"),
    (["-n", "1"], 1, 64, "cube")])
def test_cpp_tree_matches_oracle(tree_dump, argv, n, ncrit, dist):
    if not argv:  # the defaults only: parse, do not build a 1M tree here
        argv = ["-n", "2000"]
        n = 2000
    head, lo_hi, b, c = _run_dump(tree_dump, *argv)
    assert int(head[0]) == n and int(head[1]) == ncrit and head[4] == dist
    assert int(head[2]) == 10 and float(head[3]) == 0.4  # Args defaults (include/args.h)
    x = O.data_cube(n, 0) if dist == "cube" else O.data_plummer(n, 0)
    q = O.data_source(n, 0)
    lo, hi = O.bounds(x)
    assert np.array_equal(lo_hi, np.concatenate([lo, hi]))
    t = O.Tree(x, ncrit)
    assert np.array_equal(b[:, 0].astype(int), t.perm)  # IBODY keeps the input index
    assert np.array_equal(b[:, 1:4], x[t.perm]) and np.array_equal(b[:, 4], q[t.perm])
    assert len(c) == t.ncell
    for k, name in enumerate(("iparent", "ichild", "nchild", "ibody", "nbody")):
        assert np.array_equal(c[:, k].astype(int), getattr(t, name)), name
    assert np.array_equal(c[:, 5:8], t.X) and np.array_equal(c[:, 8], t.R)
    assert np.array_equal(c[:, 9].astype(int), t.ibody)  # Cell::BODY points at IBODY


def test_cpp_args_rejects_bad_distribution(tree_dump):
    r = subprocess.run([tree_dump, "-n", "10", "-d", "x"], capture_output=True, text=True)
    assert r.returncode != 0


def test_cpp_dipole_correction_matches_oracle(tmp_path):
    """UpDownPass::getDipole + dipoleCorrection (include/up_down_pass.h:77-99) on host Bodies."""
    F = _F()
    exe = str(tmp_path / "dipole_check")
    libdir = os.path.dirname(F.LIB_PATH)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "dipole_check.cpp"), "-o", exe,
                    "-L", libdir, "-lexafmm_amd", "-Wl,-rpath," + libdir], check=True)
    rng = np.random.default_rng(0)
    n = 300
    x = rng.random((n, 3)) * 2 * np.pi - np.pi
    q = rng.random(n) - 0.3
    trg = rng.random((n, 4))
    inp = f"{n}\n" + "\n".join(" ".join(repr(float(v)) for v in (*x[i], q[i], *trg[i])) for i in range(n))
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    got = np.array([l.split() for l in out.strip().split("\n")], float)
    ref = O.dipole_correction(x, q, 2 * np.pi, trg.copy())
    assert np.allclose(got, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_example_driver_prints_reference_errors():
    """examples/fmm.cxx defaults on 32768 cube bodies print
    4.7429586e-06 / 4.8780842e-06 (the reference's output)."""
    exe = os.path.join(PKG, "lib", "fmm")
    out = subprocess.run([exe, "-n", "32768", "-v"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = dict(re.findall(r"Rel\. L2 Error \((pot|acc)\)\s*:\s*(\S+)", out.stdout))
    assert got == {"pot": "4.7429586e-06", "acc": "4.8780842e-06"}, out.stdout
    assert re.search(r"Cells\s*:\s*\d+", out.stdout) and re.search(r"M2L calls\s*:\s*\d+", out.stdout)


@pytest.mark.gpu
@pytest.mark.parametrize("images", [1, 2])
def test_example_driver_periodic(images):
    """examples/fmm.cxx with --images (cycle 2 pi, the example's): the printed
    FMM-vs-direct errors equal the oracle's for the same system, 100 sampled
    targets and the same shifts.  At images = 1 both sums cover the 27 boxes;
    at images >= 2 the reference's FMM (traversePeriodic, include/traversal.h:
    233-287) and its direct sum (include/traversal.h:355-377) cover different
    sets of boxes, so the printed error is that difference, not FMM error."""
    n, P, theta, cyc = 2048, 10, 0.4, 2 * np.pi
    exe = os.path.join(PKG, "lib", "fmm")
    out = subprocess.run([exe, "-n", str(n), "-i", str(images), "-v"], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    got = {k: float(v) for k, v in re.findall(r"Rel\. L2 Error \((pot|acc)\)\s*:\s*(\S+)", out.stdout)}
    assert set(got) == {"pot", "acc"}, out.stdout

    x, q = O.data_cube(n, 0), O.data_source(n, 0)
    t = O.Tree(x, 64)
    xs, qs = x[t.perm], q[t.perm]
    k = O.Kernel(P)
    fmm = t.fmm_periodic(k, xs, qs, theta, 5000, images, cyc)
    idx = np.arange(100) * (n // 100)  # Dataset::sampleBodies (include/dataset.h)
    neighbor = int(2 * (0.5 / theta + 1) + 1)
    prange = sum(neighbor ** l for l in range(images))
    d = np.zeros((len(idx), 4))
    for ix in range(-prange, prange + 1):
        for iy in range(-prange, prange + 1):
            for iz in range(-prange, prange + 1):
                d += O.direct(k, xs[idx], xs + np.array([ix, iy, iz]) * cyc, qs)
    f = fmm[idx]
    pot = np.sqrt(np.sum((d[:, 0] - f[:, 0]) ** 2) / np.sum(d[:, 0] ** 2))
    acc = np.sqrt(np.sum((d[:, 1:] - f[:, 1:]) ** 2) / np.sum(d[:, 1:] ** 2))
    assert abs(got["pot"] - pot) <= 1e-6 * pot and abs(got["acc"] - acc) <= 1e-6 * acc, (got, pot, acc)
    if images == 1:
        assert pot < 1e-4 and acc < 1e-4


@pytest.mark.gpu
def test_example_driver_writes_list_counts(tmp_path):
    """examples/fmm.cxx's writeList after traverse: per cell the Morton key
    and the sizes of its P2P and M2L lists, equal to the oracle's."""
    exe = os.path.join(PKG, "lib", "fmm")
    path = str(tmp_path) + "/"
    out = subprocess.run([exe, "-n", "4096", "-p", path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = np.array([l.split() for l in open(path + "list000000.dat").read().strip().split("\n")], np.uint64)
    x = O.data_cube(4096, 0)
    t = O.Tree(x, 64)
    ci, cj, ty = t.traverse(0.4, 5000)
    assert np.array_equal(rows[:, 0], t.icell)
    assert np.array_equal(rows[:, 1].astype(int), np.bincount(ci[ty == 1], minlength=t.ncell))
    assert np.array_equal(rows[:, 2].astype(int), np.bincount(ci[ty == 0], minlength=t.ncell))


@pytest.mark.gpu
def test_set_tree_equals_set_bodies():
    F = _F()
    x = O.data_plummer(6000, 4)
    q = O.data_source(6000, 4)
    t = O.Tree(x, 32)
    tree = {"iparent": t.iparent, "ichild": t.ichild, "nchild": t.nchild, "ibody": t.ibody,
            "nbody": t.nbody, "X": t.X, "R": t.R}
    out = []
    for use_tree in (False, True):
        f = F.FMM(P=8, ncrit=32, m2l_mode=F.M2L_TARGET)
        if use_tree:
            f.set_tree(x[t.perm], q[t.perm], tree)
        else:
            f.set_bodies(x, q)
        f.evaluate()
        out.append((f.targets(input_order=False), f.expansions(), f.stats()))
        f.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1][0], out[1][1][0]) and np.array_equal(out[0][1][1], out[1][1][1])
    assert out[0][2]["m2l_pairs"] == out[1][2]["m2l_pairs"]


def _tree_dict(t, **over):
    d = {"iparent": t.iparent, "ichild": t.ichild, "nchild": t.nchild, "ibody": t.ibody,
         "nbody": t.nbody, "X": t.X, "R": t.R}
    d.update({k: v for k, v in over.items()})
    return d


@pytest.mark.parametrize("theta,nspawn", [(0.4, 5000), (0.3, 200)])
def test_plan_from_tree_lists_match_oracle(theta, nspawn):
    F = _F()
    x = O.data_plummer(8000, 6)
    t = O.Tree(x, 16)
    pl = F.Plan(x[t.perm], theta=theta, nspawn=nspawn, tree=_tree_dict(t))
    tr = pl.tree()
    assert np.array_equal(tr["perm"], np.arange(len(x))) and tr["depth"] == t.depth
    ci, cj, ty = t.traverse(theta, nspawn)
    mp, ms, pp, ps = pl.lists()
    for typ, ptr, srcs in ((0, mp, ms), (1, pp, ps)):
        sel = ty == typ
        a, b = ci[sel], cj[sel]
        order = np.argsort(a, kind="stable")
        optr = np.zeros(t.ncell + 1, np.int64)
        optr[1:] = np.cumsum(np.bincount(a, minlength=t.ncell))
        assert np.array_equal(ptr, optr) and np.array_equal(srcs, b[order])


def test_plan_from_tree_rejects_inconsistent_trees():
    F = _F()
    x = O.data_cube(500, 1)
    t = O.Tree(x, 64)
    xs = x[t.perm]
    leaf = int(np.nonzero(t.nchild == 0)[0][-1])
    ib = t.ibody.copy()
    ib[leaf] += 1                       # leaves no longer partition the bodies
    nb = t.nbody.copy()
    nb[0] -= 1                          # root does not cover every body
    ip = t.iparent.copy()
    ip[leaf] = leaf                     # parent link broken
    ic = t.ichild.copy()
    ic[0] = 0                           # child block does not follow the parent
    for bad in (dict(ibody=ib), dict(nbody=nb), dict(iparent=ip), dict(ichild=ic)):
        with pytest.raises(F.FMMError, match="fmm_plan_from_tree"):
            F.Plan(xs, tree=_tree_dict(t, **bad))
    F.Plan(xs, tree=_tree_dict(t))      # the unmodified tree is accepted


def test_cpp_morton_keys_and_write_list(tree_dump, tmp_path):
    """BuildTree's ICELL (BuildTree::getKey, include/build_tree.h:111-120,132)
    equals the oracle's; Traversal::writeList writes "ICELL numP2P numM2L" per
    cell to <path>list<rank>.dat (include/traversal.h:412-421)."""
    path = str(tmp_path) + "/"
    head, lo_hi, b, c = _run_dump(tree_dump, "-n", "3000", "-c", "16", "-p", path)
    out = subprocess.run([tree_dump, "-n", "3000", "-c", "16", "-p", path], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    nb, nc = map(int, out[2].split())
    keys = np.array([int(l.split()[10]) for l in out[3 + nb:3 + nb + nc]], np.uint64)
    x = O.data_cube(3000, 0)
    t = O.Tree(x, 16)
    assert np.array_equal(keys, t.icell)
    rows = [l.split() for l in open(path + "list000003.dat").read().strip().split("\n")]
    assert len(rows) == nc
    for i, r in enumerate(rows):
        assert (int(r[0]), int(r[1]), int(r[2])) == (int(keys[i]), i % 7, i % 11)


def test_plan_list_counts_match_oracle():
    """fmm_plan_list_counts: numP2P / numM2L of every cell are the oracle
    traversal's per-target pair counts; the keys nest (a child's key is its
    parent's shifted by 3 bits plus its octant, above the level offset)."""
    F = _F()
    x = O.data_plummer(20000, 3)
    pl = F.Plan(x, ncrit=32, theta=0.4, nspawn=1000)
    key, np2p, nm2l = pl.list_counts()
    t = O.Tree(x, 32)
    ci, cj, ty = t.traverse(0.4, 1000)
    assert np.array_equal(np2p, np.bincount(ci[ty == 1], minlength=t.ncell))
    assert np.array_equal(nm2l, np.bincount(ci[ty == 0], minlength=t.ncell))
    assert np.array_equal(key, t.icell)
    tr = pl.tree()
    off = lambda l: ((1 << (3 * l)) - 1) // 7
    level = np.zeros(t.ncell, int)
    for cc in range(1, t.ncell):
        level[cc] = level[tr["iparent"][cc]] + 1
    for cc in range(1, t.ncell):
        p, lc = tr["iparent"][cc], level[cc]
        oct_ = sum(1 << d for d in range(3) if tr["X"][cc][d] > tr["X"][p][d])
        assert (int(key[cc]) - off(lc)) >> 3 == int(key[p]) - off(lc - 1)
        assert (int(key[cc]) - off(lc)) & 7 == oct_


def test_kahan_compensates():
    """kahan<T> (include/kahan.h) in the C++ layer: 1 + 2^20 terms each below
    half an ulp of the sum.  The plain sum never moves; the compensated value
    is the exact sum to a few ulps (float: 1 + 2^20 * 1e-8 = 1.01048576, the
    terms being the float nearest 1e-8), and the compensation term is non-zero."""
    exe = os.path.join(os.environ.get("TMPDIR", "/tmp"), "kahan_check_%d" % os.getpid())
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "kahan_check.cpp"), "-o", exe], check=True)
    out = {l.split()[0]: [float(v) for v in l.split()[1:]]
           for l in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()}
    os.unlink(exe)
    small32 = float(np.float32(1e-8))
    plain, val, s, c = out["float"]
    exact = 1.0 + (1 << 20) * small32
    assert plain == 1.0
    assert abs(val - exact) <= 4 * np.spacing(np.float32(exact)) and c != 0.0
    plain, val, s, c = out["double"]
    exact = 1.0 + (1 << 20) * 1e-17
    assert plain == 1.0
    assert abs(val - exact) <= 4 * np.spacing(exact) and c != 0.0
    s_val, m_val, neg, b = out["ops"]
    assert abs(s_val - 1.0) <= 2 * np.spacing(1.0)
    assert abs(m_val - (1.0 + 1e-14) / 2) <= 2 * np.spacing(0.5)
    assert abs(neg + 1.0 + 1e-14) <= 2 * np.spacing(1.0)
    assert abs(b + 1e-14) <= 1e-14 * 1e-10
