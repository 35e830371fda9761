"""Fixtures written by tests/golden/make_golden.py from exafmm-beta's own
compiled operators and generators: the oracle (CPU) and the HIP path (GPU)
against them, so parity does not depend on the reference tree being present."""
import os

import numpy as np
import pytest

import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name))  # arrays only: allow_pickle stays False


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("P", [8, 10, 12])
def test_oracle_operators_match_fixture(P):
    f = _load("ops_p%d.npz" % P)
    k = O.Kernel(P)
    M = k.p2m(f["Xc"], f["xb"], f["q"])
    assert _rel(M, f["M"]) < 1e-14
    assert _rel(k.m2m(f["Xp"], f["Xc"], f["M"]), f["Mp"]) < 1e-13
    assert _rel(k.m2l(f["Xi"], f["Xp"], f["Mp"]), f["L"]) < 1e-13
    assert _rel(k.l2l(f["Xl"], f["Xi"], f["L"]), f["Ll"]) < 1e-13
    assert _rel(k.l2p(f["Xl"], f["Ll"], f["xt"]), f["l2p"]) < 1e-13
    # the reference's SIMD P2P seeds rsqrt in float and takes two Newton steps
    assert _rel(k.p2p(f["xt"], f["xb"], f["q"]), f["p2p"]) < 1e-12


def test_oracle_kernel_chain_matches_fixture():
    f = _load("chain.npz")
    for P in (4, 10, 20):
        k = O.Kernel(P)
        src = np.array([[2.0, 2.0, 2.0]])
        M = k.p2m([3, 1, 1], src, [1.0])
        MJ = k.m2m([4, 0, 0], [3, 1, 1], M)
        LI = k.m2l([-4, 0, 0], [4, 0, 0], MJ)
        Li = k.l2l([-3, 1, 1], [-4, 0, 0], LI)
        fmm = k.l2p([-3, 1, 1], Li, np.array([[-2.0, 2.0, 2.0]]))
        assert _rel(fmm, f["fmm_p%d" % P]) < 1e-12, P
        assert _rel(k.p2p(np.array([[-2.0, 2.0, 2.0]]), src, [1.0]), f["direct_p%d" % P]) < 1e-14


def test_oracle_fmm4096_matches_fixture():
    f = _load("fmm4096.npz")
    x, q = f["x"], f["q"]
    assert np.array_equal(x, O.data_cube(4096, 0)) and np.array_equal(q, O.data_source(4096, 0))
    t = O.Tree(x, 64)
    assert np.array_equal(t.perm, f["perm"]) and t.ncell == int(f["ncell"])
    ci, cj, ty = t.traverse(0.4, 5000)
    assert int((ty == 0).sum()) == int(f["m2l_pairs"]) and int((ty == 1).sum()) == int(f["p2p_pairs"])
    trg = t.fmm(O.Kernel(10), x[t.perm], q[t.perm], 0.4, 5000)
    assert _rel(trg[:, 0], f["trg"][:, 0]) < 1e-12 and _rel(trg[:, 1:], f["trg"][:, 1:]) < 1e-12


@pytest.mark.gpu
def test_gpu_fmm4096_matches_fixture():
    import exafmm_amd as F
    f = _load("fmm4096.npz")
    g = F.FMM(P=10)
    g.set_bodies(f["x"], f["q"])
    g.evaluate()
    assert np.array_equal(g.tree()["perm"], f["perm"])
    trg = g.targets(input_order=False)
    assert _rel(trg[:, 0], f["trg"][:, 0]) < 1e-10 and _rel(trg[:, 1:], f["trg"][:, 1:]) < 1e-10
    s = g.stats()
    assert (s["m2l_pairs"], s["p2p_pairs"]) == (int(f["m2l_pairs"]), int(f["p2p_pairs"]))
    g.close()
