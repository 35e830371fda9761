"""Writes the golden fixtures of this directory from exafmm-beta's own compiled
operators and generators (oracle/_ref/libexafmm_ref.so, built by
oracle/Makefile from kernels/laplace.h and include/dataset.h in place).

  python tests/golden/make_golden.py

ops_p{P}.npz    one random chain P2M -> M2M -> M2L -> L2L -> L2P and a P2P
                block through the reference operators, inputs and outputs
                (reference normalisation of M and L), P = 8, 10, 12
chain.npz       examples/kernel.cxx's unit-source chain for P = 4, 10, 20
fmm4096.npz     N = 4096 reference cube bodies (Dataset::initBodies seed 0),
                every body's (pot, ax, ay, az) after a P = 10 solve with the
                reference operators over the traversal's lists
                (theta = 0.4, ncrit = 64, nspawn = 5000), in tree order

Arrays only (numpy savez, no pickled objects).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def ops(P):
    rng = np.random.default_rng(100 + P)
    k = O.Kernel(P, use_ref=True)
    Xc = rng.uniform(-1, 1, 3)
    xb = Xc + rng.uniform(-0.5, 0.5, (40, 3))
    q = rng.uniform(-1, 1, 40)
    Xp = Xc + np.array([0.5, -0.5, 0.5])
    Xi = Xp + np.array([4.0, 1.0, -3.0])
    Xl = Xi + np.array([-0.25, 0.5, 0.25])
    xt = Xl + rng.uniform(-0.25, 0.25, (24, 3))
    M = k.p2m(Xc, xb, q)
    Mp = k.m2m(Xp, Xc, M)
    L = k.m2l(Xi, Xp, Mp)
    Ll = k.l2l(Xl, Xi, L)
    l2p = k.l2p(Xl, Ll, xt)
    p2p = k.p2p(xt, xb, q)
    return dict(Xc=Xc, xb=xb, q=q, Xp=Xp, Xi=Xi, Xl=Xl, xt=xt, M=M, Mp=Mp, L=L, Ll=Ll, l2p=l2p, p2p=p2p)


def chain():
    out = {}
    for P in (4, 10, 20):
        k = O.Kernel(P, use_ref=True)
        src = np.array([[2.0, 2.0, 2.0]])
        M = k.p2m([3, 1, 1], src, [1.0])
        MJ = k.m2m([4, 0, 0], [3, 1, 1], M)
        LI = k.m2l([-4, 0, 0], [4, 0, 0], MJ)
        Li = k.l2l([-3, 1, 1], [-4, 0, 0], LI)
        out["fmm_p%d" % P] = k.l2p([-3, 1, 1], Li, np.array([[-2.0, 2.0, 2.0]]))
        out["direct_p%d" % P] = k.p2p(np.array([[-2.0, 2.0, 2.0]]), src, [1.0])
    return out


def fmm4096():
    x, q = O.ref_init_bodies(4096, "cube", 0)
    t = O.Tree(x, 64)
    xs, qs = x[t.perm], q[t.perm]
    ci, cj, ty = t.traverse(0.4, 5000)
    trg = t.ref_eval(O.Kernel(10, use_ref=True), xs, qs, ci, cj, ty)
    return dict(x=x, q=q, perm=t.perm, trg=trg, ncell=np.int64(t.ncell),
                m2l_pairs=np.int64((ty == 0).sum()), p2p_pairs=np.int64((ty == 1).sum()))


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (needs the reference sources): make -C oracle ref")
    for P in (8, 10, 12):
        np.savez(os.path.join(HERE, "ops_p%d.npz" % P), **ops(P))
    np.savez(os.path.join(HERE, "chain.npz"), **chain())
    np.savez(os.path.join(HERE, "fmm4096.npz"), **fmm4096())
    print("written to", HERE)


if __name__ == "__main__":
    main()
