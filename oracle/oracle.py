"""ctypes view of the oracle (oracle.c) and of the reference-operator harness
(ref_harness.cpp, built into oracle/_ref when the exafmm-beta tree is present).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.  Arrays are numpy; complex
expansions are numpy complex128 of shape (ncell, NTERM).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libexafmm_ref.so")
# the same reference sources built without -ffast-math (IEEE libm calls and no
# reassociation): used to show that the generators differ only by that flag
REF_IEEE_SO = os.path.join(HERE, "_ref", "libexafmm_ref_ieee.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build(quiet=True):
    """Compile the oracle (and _ref when the reference tree is present)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    return ctypes.CDLL(path)


_lib = None
_ref = None
_ref_ieee = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = _load(ORACLE_SO)
        L.orc_kernel_create.restype = _vp
        L.orc_kernel_create.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_kernel_destroy.argtypes = [_vp]
        L.orc_kernel_nterm.argtypes = [_vp]
        L.orc_p2p.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.orc_p2m.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.orc_m2m.argtypes = [_vp, _dp, _dp, _dp, _dp]
        L.orc_m2l.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
        L.orc_l2l.argtypes = [_vp, _dp, _dp, _dp, _dp]
        L.orc_l2p.argtypes = [_vp, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_bounds.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_bounds_update.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_tree_build.restype = _vp
        L.orc_tree_build.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.orc_tree_destroy.argtypes = [_vp]
        L.orc_tree_ncell.argtypes = [_vp]
        L.orc_tree_depth.argtypes = [_vp]
        L.orc_tree_get.argtypes = [_vp, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp,
                                   ctypes.POINTER(ctypes.c_uint64)]
        L.orc_traverse.restype = _vp
        L.orc_traverse.argtypes = [_vp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_traverse_pair.restype = _vp
        L.orc_traverse_pair.argtypes = [_vp, _vp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_fmm_pair.argtypes = [_vp, _vp, _dp, _vp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_lists_destroy.argtypes = [_vp]
        L.orc_lists_count.restype = ctypes.c_int64
        L.orc_lists_count.argtypes = [_vp, ctypes.c_int]
        L.orc_lists_get.argtypes = [_vp, _ip, _ip, _ip]
        L.orc_fmm.argtypes = [_vp, _vp, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.orc_fmm_periodic.argtypes = [_vp, _vp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                       _dp, _dp]
        L.orc_wrap.argtypes = [_dp, _dp]
        L.orc_wrap.restype = ctypes.c_int
        L.orc_dipole_correction.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_ewald.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, _dp, _dp]
        L.orc_eval_lists.argtypes = [_vp, _vp, _dp, _dp, ctypes.c_int64, _ip, _ip, _ip, _dp,
                                     _dp, _dp]
        L.orc_upward.argtypes = [_vp, _vp, _dp, _dp, _dp]
        L.orc_downward.argtypes = [_vp, _vp, _dp, _dp, _dp]
        L.orc_direct.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp]
        L.orc_rel_l2.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_data_cube.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_data_plummer.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_data_source.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_sample_indices.argtypes = [ctypes.c_int, ctypes.c_int, _ip]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = _load(REF_SO)
        R.ref_kernel_create.restype = _vp
        R.ref_kernel_create.argtypes = [ctypes.c_int, ctypes.c_double]
        R.ref_kernel_destroy.argtypes = [_vp]
        R.ref_p2p.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]
        R.ref_p2m.argtypes = [_vp, _dp, ctypes.c_int, _dp, _dp, _dp]
        R.ref_m2m.argtypes = [_vp, _dp, _dp, _dp, _dp]
        R.ref_m2l.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
        R.ref_l2l.argtypes = [_vp, _dp, _dp, _dp, _dp]
        R.ref_l2p.argtypes = [_vp, _dp, _dp, ctypes.c_int, _dp, _dp]
        R.ref_eval.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip,
                               _dp, _dp, ctypes.c_long, _ip, _ip, _ip, _dp, _dp, _dp]
        R.ref_init_bodies.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _dp, _dp]
        R.ref_time.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip,
                               _dp, _dp, ctypes.c_long, _ip, _ip, ctypes.c_long, _ip, _ip, _dp, _dp]
        R.ref_time_sampled.argtypes = [_vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _ip,
                                       _ip, _dp, _dp, ctypes.c_int, ctypes.c_long, _ip, _ip,
                                       ctypes.c_long, _ip, _ip, _dp, _dp]
        _ref = R
    return _ref


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


class Kernel:
    """Oracle Kernel(P): the six operators of kernels/laplace.h (or, with
    use_ref=True, the reference's own compiled operators)."""

    def __init__(self, P, eps2=0.0, use_ref=False):
        self.P = P
        self.nterm = P * (P + 1) // 2
        self.use_ref = use_ref
        if use_ref:
            self._L = ref()
            self.h = self._L.ref_kernel_create(P, eps2)
            self._pre = "ref_"
        else:
            self._L = lib()
            self.h = self._L.orc_kernel_create(P, eps2)
            self._pre = "orc_"

    def __del__(self):
        try:
            getattr(self._L, self._pre + "kernel_destroy")(self.h)
        except Exception:
            pass

    def _f(self, name):
        return getattr(self._L, self._pre + name)

    def p2p(self, xi, xj, qj, xper=None, trg=None):
        xi, xj, qj = _c(xi), _c(xj), _c(qj)
        trg = np.zeros((len(xi), 4)) if trg is None else trg
        xp = _c(xper if xper is not None else np.zeros(3))
        self._f("p2p")(self.h, len(xi), _ptr(xi), _ptr(trg), len(xj), _ptr(xj), _ptr(qj), _ptr(xp))
        return trg

    def p2m(self, Xc, x, q, M=None):
        M = np.zeros(self.nterm, np.complex128) if M is None else M
        x, q = _c(x), _c(q)
        self._f("p2m")(self.h, _ptr(_c(Xc)), len(x), _ptr(x), _ptr(q), _ptr(M.view(np.float64)))
        return M

    def m2m(self, Xp, Xc, Mc, Mp=None):
        Mp = np.zeros(self.nterm, np.complex128) if Mp is None else Mp
        Mc = _c(Mc, np.complex128)
        self._f("m2m")(self.h, _ptr(_c(Xp)), _ptr(_c(Xc)), _ptr(Mc.view(np.float64)),
                       _ptr(Mp.view(np.float64)))
        return Mp

    def m2l(self, Xi, Xj, Mj, Li=None, xper=None):
        Li = np.zeros(self.nterm, np.complex128) if Li is None else Li
        Mj = _c(Mj, np.complex128)
        xp = _c(xper if xper is not None else np.zeros(3))
        self._f("m2l")(self.h, _ptr(_c(Xi)), _ptr(_c(Xj)), _ptr(xp), _ptr(Mj.view(np.float64)),
                       _ptr(Li.view(np.float64)))
        return Li

    def l2l(self, Xc, Xp, Lp, Lc=None):
        Lc = np.zeros(self.nterm, np.complex128) if Lc is None else Lc
        Lp = _c(Lp, np.complex128)
        self._f("l2l")(self.h, _ptr(_c(Xc)), _ptr(_c(Xp)), _ptr(Lp.view(np.float64)),
                       _ptr(Lc.view(np.float64)))
        return Lc

    def l2p(self, Xc, L, x, trg=None):
        x = _c(x)
        trg = np.zeros((len(x), 4)) if trg is None else trg
        L = _c(L, np.complex128)
        self._f("l2p")(self.h, _ptr(_c(Xc)), _ptr(L.view(np.float64)), len(x), _ptr(x), _ptr(trg))
        return trg


def bounds(x, start=None):
    """BoundBox::getBounds(bodies), or getBounds(bodies, start) widening the
    given (lo, hi) (include/bound_box.h:11-37)."""
    x = _c(x)
    if start is None:
        lo, hi = np.zeros(3), np.zeros(3)
        lib().orc_bounds(len(x), _ptr(x), _ptr(lo), _ptr(hi))
    else:
        lo, hi = _c(start[0]).copy(), _c(start[1]).copy()
        lib().orc_bounds_update(len(x), _ptr(x), _ptr(lo), _ptr(hi))
    return lo, hi


def ine_j_bodies(n, seed_i=0, seed_j=1, dist="cube"):
    """examples/fmm.cxx:37-47 with -j: targets of seed 0 with x -> (x + pi) / 2,
    sources of seed 1 with x -> (x - pi) / 2 (for the cube: targets in
    [0, pi] and sources in [-pi, 0] along x), charges of each set by its own seed."""
    gen = data_cube if dist == "cube" else data_plummer
    xi, xj = gen(n, seed_i), gen(n, seed_j)
    xi[:, 0] = (xi[:, 0] + np.pi) * 0.5
    xj[:, 0] = (xj[:, 0] - np.pi) * 0.5
    return xi, data_source(n, seed_i), xj, data_source(n, seed_j)


class Tree:
    """BuildTree::buildTree output: perm (tree order -> input index) and the
    CellBase fields."""

    def __init__(self, x, ncrit=64, bnds=None):
        L = lib()
        x = _c(x)
        self.n = len(x)
        lo, hi = bnds if bnds is not None else bounds(x)
        self.h = L.orc_tree_build(self.n, _ptr(x), ncrit, _ptr(_c(lo)), _ptr(_c(hi)))
        nc = L.orc_tree_ncell(self.h)
        self.ncell = nc
        self.depth = L.orc_tree_depth(self.h)
        self.perm = np.zeros(self.n, np.int32)
        self.iparent, self.ichild, self.nchild, self.ibody, self.nbody = (
            np.zeros(nc, np.int32) for _ in range(5))
        self.X = np.zeros((nc, 3))
        self.R = np.zeros(nc)
        self.icell = np.zeros(nc, np.uint64)
        L.orc_tree_get(self.h, _ptr(self.perm, _ip), _ptr(self.iparent, _ip),
                       _ptr(self.ichild, _ip), _ptr(self.nchild, _ip), _ptr(self.ibody, _ip),
                       _ptr(self.nbody, _ip), _ptr(self.X), _ptr(self.R),
                       self.icell.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))

    def __del__(self):
        try:
            lib().orc_tree_destroy(self.h)
        except Exception:
            pass

    def traverse(self, theta=0.4, nspawn=5000, xper=None):
        L = lib()
        xp = _c(xper if xper is not None else np.zeros(3))
        h = L.orc_traverse(self.h, theta, nspawn, _ptr(xp))
        n = L.orc_lists_count(h, -1)
        ci, cj, ty = (np.zeros(n, np.int32) for _ in range(3))
        L.orc_lists_get(h, _ptr(ci, _ip), _ptr(cj, _ip), _ptr(ty, _ip))
        L.orc_lists_destroy(h)
        return ci, cj, ty

    def traverse_pair(self, sources, theta=0.4, nspawn=5000, xper=None):
        """Traversal::traverse(icells, jcells): this tree's cells as targets,
        `sources` (another Tree) as sources; ci indexes this tree, cj the other."""
        L = lib()
        xp = _c(xper if xper is not None else np.zeros(3))
        h = L.orc_traverse_pair(self.h, sources.h, theta, nspawn, _ptr(xp))
        n = L.orc_lists_count(h, -1)
        ci, cj, ty = (np.zeros(n, np.int32) for _ in range(3))
        L.orc_lists_get(h, _ptr(ci, _ip), _ptr(cj, _ip), _ptr(ty, _ip))
        L.orc_lists_destroy(h)
        return ci, cj, ty

    def fmm_pair(self, kernel, x_sorted, sources, xj_sorted, qj_sorted, theta=0.4, nspawn=5000):
        """orc_fmm_pair: the field of the source tree's bodies at this tree's bodies."""
        xs, xj, qj = _c(x_sorted), _c(xj_sorted), _c(qj_sorted)
        trg = np.zeros((self.n, 4))
        lib().orc_fmm_pair(kernel.h, self.h, _ptr(xs), sources.h, _ptr(xj), _ptr(qj), theta, nspawn,
                           _ptr(trg))
        return trg

    def fmm(self, kernel, x_sorted, q_sorted, theta=0.4, nspawn=5000, want_ml=False):
        L = lib()
        xs, qs = _c(x_sorted), _c(q_sorted)
        trg = np.zeros((self.n, 4))
        M = Lc = None
        if want_ml:
            M = np.zeros((self.ncell, kernel.nterm), np.complex128)
            Lc = np.zeros((self.ncell, kernel.nterm), np.complex128)
        L.orc_fmm(kernel.h, self.h, _ptr(xs), _ptr(qs), theta, nspawn, _ptr(trg),
                  _ptr(M.view(np.float64)) if want_ml else None,
                  _ptr(Lc.view(np.float64)) if want_ml else None)
        return (trg, M, Lc) if want_ml else trg

    def fmm_periodic(self, kernel, x_sorted, q_sorted, theta, nspawn, images, cycle):
        """orc_fmm_periodic: 27 image traversals + traversePeriodic (include/traversal.h)."""
        xs, qs = _c(x_sorted), _c(q_sorted)
        trg = np.zeros((self.n, 4))
        cyc = _c(np.broadcast_to(np.asarray(cycle, dtype=np.float64), (3,)))
        lib().orc_fmm_periodic(kernel.h, self.h, _ptr(xs), _ptr(qs), theta, nspawn, images, _ptr(cyc),
                               _ptr(trg))
        return trg

    def upward(self, kernel, x_sorted, q_sorted):
        M = np.zeros((self.ncell, kernel.nterm), np.complex128)
        xs, qs = _c(x_sorted), _c(q_sorted)
        lib().orc_upward(kernel.h, self.h, _ptr(xs), _ptr(qs), _ptr(M.view(np.float64)))
        return M

    def eval_lists(self, kernel, x_sorted, q_sorted, ci, cj, ty, M, trg=None, Lc=None):
        xs, qs = _c(x_sorted), _c(q_sorted)
        trg = np.zeros((self.n, 4)) if trg is None else trg
        Lc = np.zeros((self.ncell, kernel.nterm), np.complex128) if Lc is None else Lc
        ci, cj, ty = (_c(a, np.int32) for a in (ci, cj, ty))
        M = _c(M, np.complex128)
        lib().orc_eval_lists(kernel.h, self.h, _ptr(xs), _ptr(qs), len(ci), _ptr(ci, _ip),
                             _ptr(cj, _ip), _ptr(ty, _ip), _ptr(trg), _ptr(Lc.view(np.float64)),
                             _ptr(M.view(np.float64)))
        return trg, Lc

    def downward(self, kernel, x_sorted, Lc, trg):
        xs = _c(x_sorted)
        lib().orc_downward(kernel.h, self.h, _ptr(xs), _ptr(Lc.view(np.float64)), _ptr(trg))
        return trg

    def ref_eval(self, ref_kernel, x_sorted, q_sorted, ci, cj, ty, want_ml=False):
        """The reference's own operators over this tree and pair list."""
        R = ref()
        xs, qs = _c(x_sorted), _c(q_sorted)
        trg = np.zeros((self.n, 4))
        nt = ref_kernel.nterm
        M = np.zeros((self.ncell, nt), np.complex128)
        Lc = np.zeros((self.ncell, nt), np.complex128)
        ci, cj, ty = (_c(a, np.int32) for a in (ci, cj, ty))
        R.ref_eval(ref_kernel.h, self.n, _ptr(xs), _ptr(qs), self.ncell, _ptr(self.iparent, _ip),
                   _ptr(self.ichild, _ip), _ptr(self.nchild, _ip), _ptr(self.ibody, _ip),
                   _ptr(self.nbody, _ip), _ptr(_c(self.X)), _ptr(self.R), len(ci), _ptr(ci, _ip),
                   _ptr(cj, _ip), _ptr(ty, _ip), _ptr(trg), _ptr(M.view(np.float64)),
                   _ptr(Lc.view(np.float64)))
        return (trg, M, Lc) if want_ml else trg


def ref_time(tree, ref_kernel, x_sorted, q_sorted, m2l_pairs, p2p_pairs):
    """Time the reference operators (upward, given M2L pairs, given P2P pairs,
    downward) on one core; returns seconds (upward, m2l, p2p, downward)."""
    R = ref()
    t = tree
    xs, qs = _c(x_sorted), _c(q_sorted)
    trg = np.zeros((t.n, 4))
    mi, mj = (_c(a, np.int32) for a in m2l_pairs)
    pi, pj = (_c(a, np.int32) for a in p2p_pairs)
    times = np.zeros(4)
    R.ref_time(ref_kernel.h, t.n, _ptr(xs), _ptr(qs), t.ncell, _ptr(t.iparent, _ip),
               _ptr(t.ichild, _ip), _ptr(t.nchild, _ip), _ptr(t.ibody, _ip), _ptr(t.nbody, _ip),
               _ptr(_c(t.X)), _ptr(t.R), len(mi), _ptr(mi, _ip), _ptr(mj, _ip), len(pi),
               _ptr(pi, _ip), _ptr(pj, _ip), _ptr(trg), _ptr(times))
    return tuple(times)


def ref_time_sampled(ref_kernel, tree, x_sorted, q_sorted, stride, m2l_pairs, p2p_pairs):
    """Time the reference operators on one core over every `stride`-th cell
    of each pass and the given M2L / P2P pairs.  tree: dict of the CellBase
    arrays (iparent, ichild, nchild, ibody, nbody, X, R).  Returns (seconds,
    work) for P2M, M2M, M2L, P2P, L2L, L2P."""
    R = ref()
    xs, qs = _c(x_sorted), _c(q_sorted)
    a = {k: _c(tree[k], np.int32) for k in ("iparent", "ichild", "nchild", "ibody", "nbody")}
    mi, mj = (_c(v, np.int32) for v in m2l_pairs)
    pi, pj = (_c(v, np.int32) for v in p2p_pairs)
    times, work = np.zeros(6), np.zeros(6)
    R.ref_time_sampled(ref_kernel.h, len(xs), _ptr(xs), _ptr(qs), len(a["nbody"]), _ptr(a["iparent"], _ip),
                       _ptr(a["ichild"], _ip), _ptr(a["nchild"], _ip), _ptr(a["ibody"], _ip),
                       _ptr(a["nbody"], _ip), _ptr(_c(tree["X"])), _ptr(_c(tree["R"])), int(stride),
                       len(mi), _ptr(mi, _ip), _ptr(mj, _ip), len(pi), _ptr(pi, _ip), _ptr(pj, _ip),
                       _ptr(times), _ptr(work))
    return times, work


def direct(kernel, xi, xj, qj):
    xi, xj, qj = _c(xi), _c(xj), _c(qj)
    trg = np.zeros((len(xi), 4))
    lib().orc_direct(kernel.h, len(xi), _ptr(xi), _ptr(trg), len(xj), _ptr(xj), _ptr(qj))
    return trg


def wrap(x, cycle):
    """vec.h wrap of each row of x into [-cycle/2, cycle/2] (a copy) and the moved-axes masks."""
    x = np.array(x, dtype=np.float64, order="C").reshape(-1, 3)
    cyc = _c(np.broadcast_to(np.asarray(cycle, dtype=np.float64), (3,)))
    iw = np.zeros(len(x), np.int32)
    for i in range(len(x)):
        row = x[i]
        buf = (ctypes.c_double * 3)(*row)
        iw[i] = lib().orc_wrap(buf, _ptr(cyc))
        x[i] = buf[:]
    return x, iw


def dipole_correction(x, q, cycle, trg):
    """UpDownPass::getDipole + dipoleCorrection (include/up_down_pass.h:77-99), in place."""
    x, q = _c(x), _c(q)
    cyc = _c(np.broadcast_to(np.asarray(cycle, dtype=np.float64), (3,)))
    assert trg.flags.c_contiguous and trg.dtype == np.float64
    lib().orc_dipole_correction(len(q), _ptr(x), _ptr(q), _ptr(cyc), _ptr(trg))
    return trg


def ewald(x, q, ksize, alpha, sigma, cutoff, cycle):
    """Ewald wave + real + self parts (include/ewald.h), FP64; returns (n, 4)."""
    x, q = _c(x), _c(q)
    cyc = _c(np.broadcast_to(np.asarray(cycle, dtype=np.float64), (3,)))
    trg = np.zeros((len(q), 4))
    lib().orc_ewald(len(q), _ptr(x), _ptr(q), ksize, alpha, sigma, cutoff, _ptr(cyc), _ptr(trg))
    return trg


def rel_l2(trg_ref, trg):
    a, b = _c(trg_ref), _c(trg)
    p, q = ctypes.c_double(), ctypes.c_double()
    lib().orc_rel_l2(len(a), _ptr(a), _ptr(b), ctypes.byref(p), ctypes.byref(q))
    return p.value, q.value


def data_cube(n, seed=0):
    x = np.zeros((n, 3))
    lib().orc_data_cube(n, seed, _ptr(x))
    return x


def data_plummer(n, seed=0):
    x = np.zeros((n, 3))
    lib().orc_data_plummer(n, seed, _ptr(x))
    return x


def data_source(n, seed=0):
    q = np.zeros(n)
    lib().orc_data_source(n, seed, _ptr(q))
    return q


def sample_indices(n, ntarget):
    idx = np.zeros(max(n, 1), np.int32)
    k = lib().orc_sample_indices(n, ntarget, _ptr(idx, _ip))
    return idx[:k]


def ref_ieee_available():
    return os.path.exists(REF_IEEE_SO)


def ref_init_bodies(n, dist="cube", seed=0, ieee=False):
    """Dataset::initBodies of the compiled reference (include/dataset.h:230-255);
    ieee=True: the build without -ffast-math."""
    global _ref_ieee
    if ieee:
        if _ref_ieee is None:
            _ref_ieee = _load(REF_IEEE_SO)
            _ref_ieee.ref_init_bodies.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _dp, _dp]
        R = _ref_ieee
    else:
        R = ref()
    x = np.zeros((n, 3))
    q = np.zeros(n)
    R.ref_init_bodies(n, dist.encode(), seed, _ptr(x), _ptr(q))
    return x, q
