"""Python binding of libexafmm_amd.so (C ABI: include/exafmm_amd.h).

Mirrors the call surface of exafmm-beta's Laplace path (examples/fmm.cxx:61-77):
``FMM(P, theta, ncrit, nspawn)`` holds what Kernel/Traversal/BuildTree hold,
``set_bodies`` does getBounds + buildTree + the traversal's list generation,
``upward``/``m2l``/``p2p``/``downward`` are the device phases and ``targets``
returns Body::TRG (potential, gradient).  There is no CPU fallback: without
the built library (or without a HIP device) every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libexafmm_amd.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)

# public symbols of include/exafmm_amd.h
SYMBOLS = [
    "fmm_default_params", "fmm_create", "fmm_destroy", "fmm_last_error", "fmm_set_bodies",
    "fmm_set_charges", "fmm_upward", "fmm_p2p", "fmm_m2l", "fmm_downward", "fmm_evaluate",
    "fmm_synchronize", "fmm_get_targets", "fmm_get_expansions", "fmm_get_stats", "fmm_get_tree",
    "fmm_get_lists", "fmm_set_bodies_plan", "fmm_set_bodies_device", "fmm_plan_build", "fmm_plan_destroy", "fmm_plan_stats",
    "fmm_plan_tree", "fmm_plan_lists", "fmm_make_bodies",
    "fmm_direct", "fmm_set_tree", "fmm_plan_from_tree", "fmm_plan_build_groups", "fmm_plan_merge",
    "fmm_server_default_config", "fmm_server_create", "fmm_server_submit", "fmm_server_wait",
    "fmm_server_get_stats", "fmm_server_destroy", "fmm_plan_prepare",
    "fmm_dipole_correction", "fmm_md_evaluate", "fmm_plan_list_counts", "fmm_get_list_counts", "fmm_ewald",
    "fmm_plan_build_tree", "fmm_plan_pair", "fmm_kernel_p2p", "fmm_kernel_p2m", "fmm_kernel_m2m",
    "fmm_kernel_m2l", "fmm_kernel_l2l", "fmm_kernel_l2p", "fmm_ewald_noalloc", "fmm_evaluate_graph",
    "fmm_plan_build_device",
]

FMM_FP64 = 0
FMM_MIXED = 1
FMM_MIXED16 = 2  # mixed with the M2L on split-FP16 operands (hi + lo halves, FP32 accumulate)
M2L_AUTO = 0      # offset groups on MFMA in target tiles (P <= 12), else per target
M2L_TARGET = 1    # per-target kernel on vector ALUs, fixed summation order
M2L_GROUPED = 2   # offset groups on MFMA, one block per group chunk, FP64 atomics
M2L_TILES = 3     # offset groups on MFMA, one block per target tile, LDS accumulation


class ServerConfig(ctypes.Structure):
    _fields_ = [("host_threads", ctypes.c_int), ("slots", ctypes.c_int),
                ("device_budget", ctypes.c_int64), ("batch_bodies", ctypes.c_int),
                ("batch_requests", ctypes.c_int), ("batch_delay_ms", ctypes.c_double),
                ("plan_on_device", ctypes.c_int)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("batch_requests", ctypes.c_int), ("batch_bodies", ctypes.c_int),
                ("slot", ctypes.c_int), ("t_submit_ms", ctypes.c_double),
                ("t_planned_ms", ctypes.c_double), ("t_started_ms", ctypes.c_double),
                ("t_done_ms", ctypes.c_double), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [("requests_submitted", ctypes.c_int64), ("requests_done", ctypes.c_int64),
                ("batches", ctypes.c_int64), ("bodies_done", ctypes.c_int64),
                ("max_batch_requests", ctypes.c_int), ("device_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FMMError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("P", ctypes.c_int), ("ncrit", ctypes.c_int), ("nspawn", ctypes.c_int),
                ("images", ctypes.c_int), ("theta", ctypes.c_double), ("eps2", ctypes.c_double),
                ("precision", ctypes.c_int), ("device", ctypes.c_int), ("m2l_mode", ctypes.c_int),
                ("cycle", ctypes.c_double * 3)]


class Stats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("ncell", ctypes.c_int), ("nleaf", ctypes.c_int),
                ("depth", ctypes.c_int), ("m2l_pairs", ctypes.c_int64),
                ("p2p_pairs", ctypes.c_int64), ("p2p_body_pairs", ctypes.c_int64),
                ("xmin", ctypes.c_double * 3), ("xmax", ctypes.c_double * 3),
                ("host_bounds_ms", ctypes.c_double), ("host_tree_ms", ctypes.c_double),
                ("host_lists_ms", ctypes.c_double), ("host_upload_ms", ctypes.c_double),
                ("dev_upward_ms", ctypes.c_double), ("dev_p2p_ms", ctypes.c_double),
                ("dev_m2l_ms", ctypes.c_double), ("dev_downward_ms", ctypes.c_double),
                ("dev_total_ms", ctypes.c_double), ("dev_m2l_kernel_ms", ctypes.c_double),
                ("dev_p2p_kernel_ms", ctypes.c_double), ("m2l_groups", ctypes.c_int64),
                ("m2l_ungrouped", ctypes.c_int64), ("m2l_path", ctypes.c_int),
                ("plan_groups_ms", ctypes.c_double)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name in ("xmin", "xmax") else v
        return d


_lib = None


def load():
    """Load the library (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FMMError("libexafmm_amd.so is not built (%s): run __graft_entry__.build()"
                           % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.fmm_default_params.argtypes = [ctypes.POINTER(Params)]
        L.fmm_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params)]
        L.fmm_destroy.argtypes = [vp]
        L.fmm_last_error.restype = ctypes.c_char_p
        L.fmm_last_error.argtypes = [vp]
        L.fmm_set_bodies.argtypes = [vp, ctypes.c_int, _dp, _dp]
        L.fmm_set_charges.argtypes = [vp, _dp]
        for f in ("fmm_upward", "fmm_p2p", "fmm_m2l", "fmm_downward", "fmm_evaluate", "fmm_evaluate_graph",
                  "fmm_synchronize", "fmm_dipole_correction"):
            getattr(L, f).argtypes = [vp]
        L.fmm_get_targets.argtypes = [vp, _dp, ctypes.c_int]
        L.fmm_get_expansions.argtypes = [vp, _dp, _dp]
        L.fmm_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.fmm_get_tree.argtypes = [vp, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp]
        L.fmm_get_lists.argtypes = [vp, _ip, _ip, _ip, _ip]
        L.fmm_set_bodies_plan.argtypes = [vp, vp, _dp, _dp]
        L.fmm_set_bodies_device.argtypes = [vp, ctypes.c_int, _dp, _dp]
        L.fmm_plan_build.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _dp, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int]
        L.fmm_plan_build_device.argtypes = [vp, ctypes.POINTER(vp), ctypes.c_int, _dp, ctypes.c_int,
                                            ctypes.c_double, ctypes.c_int]
        L.fmm_plan_destroy.argtypes = [vp]
        L.fmm_plan_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.fmm_plan_tree.argtypes = [vp, _ip, _ip, _ip, _ip, _ip, _ip, _dp, _dp]
        L.fmm_plan_lists.argtypes = [vp, _ip, _ip, _ip, _ip]
        L.fmm_make_bodies.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _dp, _dp]
        L.fmm_direct.argtypes = [vp, ctypes.c_int, _dp, _dp]
        L.fmm_md_evaluate.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.fmm_ewald.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, _dp, _dp]
        L.fmm_ewald_noalloc.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, vp, vp]
        _up = ctypes.POINTER(ctypes.c_uint64)
        L.fmm_plan_list_counts.argtypes = [vp, _up, _ip, _ip]
        L.fmm_get_list_counts.argtypes = [vp, _up, _ip, _ip]
        L.fmm_set_tree.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _ip, _ip, _ip, _ip,
                                   _dp, _dp]
        L.fmm_plan_from_tree.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _dp, ctypes.c_int, _ip, _ip,
                                         _ip, _ip, _ip, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_char_p, ctypes.c_int]
        L.fmm_plan_build_groups.argtypes = [vp, ctypes.c_int]
        L.fmm_plan_prepare.argtypes = [vp, ctypes.POINTER(Params)]
        L.fmm_server_default_config.argtypes = [ctypes.POINTER(ServerConfig)]
        L.fmm_server_default_config.restype = None
        L.fmm_server_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params),
                                        ctypes.POINTER(ServerConfig)]
        L.fmm_server_submit.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int64)]
        L.fmm_server_wait.argtypes = [vp, ctypes.c_int64, _dp, ctypes.POINTER(RequestInfo)]
        L.fmm_server_get_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.fmm_server_destroy.argtypes = [vp]
        L.fmm_server_destroy.restype = None
        L.fmm_plan_merge.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
        L.fmm_plan_build_tree.argtypes = [ctypes.POINTER(vp), ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.fmm_plan_pair.argtypes = [ctypes.POINTER(vp), vp, vp, ctypes.c_double, ctypes.c_int]
        L.fmm_kernel_p2p.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.fmm_kernel_p2m.argtypes = [vp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.fmm_kernel_m2m.argtypes = [vp, _dp, ctypes.c_int, _dp, _dp, _dp]
        L.fmm_kernel_m2l.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.fmm_kernel_l2l.argtypes = [vp, _dp, _dp, _dp, _dp]
        L.fmm_kernel_l2p.argtypes = [vp, _dp, _dp, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def make_bodies(n, dist="cube", seed=0):
    """Dataset::initBodies equivalent (host only): returns x (n,3), q (n,)."""
    x = np.zeros((n, 3))
    q = np.zeros(n)
    rc = load().fmm_make_bodies(n, dist.encode(), seed, _ptr(x), _ptr(q))
    if rc != 0:
        raise FMMError("fmm_make_bodies: unknown distribution %r (rc=%d)" % (dist, rc))
    return x, q


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t) if a is not None else None


class FMM:
    """One evaluator context (HIP stream + device arena)."""

    def __init__(self, P=10, theta=0.4, ncrit=64, nspawn=5000, eps2=0.0, device=0,
                 precision=FMM_FP64, m2l_mode=M2L_AUTO, images=0, cycle=None):
        """images > 0: periodic with period `cycle` (3 floats, or one for a cube)."""
        L = load()
        p = Params()
        L.fmm_default_params(ctypes.byref(p))
        p.P, p.theta, p.ncrit, p.nspawn, p.eps2 = P, theta, ncrit, nspawn, eps2
        p.device, p.precision, p.m2l_mode, p.images = device, precision, m2l_mode, images
        if cycle is not None:
            cyc = np.broadcast_to(np.asarray(cycle, dtype=np.float64), (3,))
            for d in range(3):
                p.cycle[d] = cyc[d]
        self.P = P
        self.nterm = P * (P + 1) // 2
        self._h = ctypes.c_void_p()
        rc = L.fmm_create(ctypes.byref(self._h), ctypes.byref(p))
        if rc != 0:
            msg = L.fmm_last_error(self._h).decode() if self._h else "fmm_create failed"
            if self._h:
                L.fmm_destroy(self._h)
                self._h = ctypes.c_void_p()
            raise FMMError("fmm_create: %s (rc=%d)" % (msg, rc))
        self.n = 0

    def close(self):
        if self._h:
            load().fmm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise FMMError("%s: %s (rc=%d)" % (what, load().fmm_last_error(self._h).decode(), rc))

    def set_bodies_plan(self, plan, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        self.n = len(plan.x)
        self._check(load().fmm_set_bodies_plan(self._h, plan._h, _ptr(plan.x), _ptr(q)),
                    "fmm_set_bodies_plan")

    def set_bodies_device(self, x, q):
        """fmm_set_bodies_device: tree, lists (and tile groups) built on the
        device, uploaded without an intermediate host copy."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if len(q) != len(x):
            raise ValueError("x and q lengths differ")
        self.n = len(x)
        self._check(load().fmm_set_bodies_device(self._h, self.n, _ptr(x), _ptr(q)), "fmm_set_bodies_device")

    def set_bodies(self, x, q):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        if len(q) != len(x):
            raise ValueError("x and q lengths differ")
        self.n = len(x)
        self._check(load().fmm_set_bodies(self._h, self.n, _ptr(x), _ptr(q)), "fmm_set_bodies")

    def set_tree(self, x_sorted, q_sorted, tree):
        """Caller-built tree (reference cell order); bodies already in tree order.
        tree: dict with iparent, ichild, nchild, ibody, nbody (int), X (ncell,3), R."""
        x = np.ascontiguousarray(x_sorted, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q_sorted, dtype=np.float64).reshape(-1)
        a = {k: np.ascontiguousarray(tree[k], dtype=np.int32)
             for k in ("iparent", "ichild", "nchild", "ibody", "nbody")}
        X = np.ascontiguousarray(tree["X"], dtype=np.float64).reshape(-1, 3)
        R = np.ascontiguousarray(tree["R"], dtype=np.float64).reshape(-1)
        self.n = len(x)
        self._check(load().fmm_set_tree(self._h, len(x), _ptr(x), _ptr(q), len(R),
                                        _ptr(a["iparent"], _ip), _ptr(a["ichild"], _ip),
                                        _ptr(a["nchild"], _ip), _ptr(a["ibody"], _ip),
                                        _ptr(a["nbody"], _ip), _ptr(X), _ptr(R)), "fmm_set_tree")

    def set_charges(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        self._check(load().fmm_set_charges(self._h, _ptr(q)), "fmm_set_charges")

    def upward(self):
        self._check(load().fmm_upward(self._h), "fmm_upward")

    def p2p(self):
        self._check(load().fmm_p2p(self._h), "fmm_p2p")

    def m2l(self):
        self._check(load().fmm_m2l(self._h), "fmm_m2l")

    def downward(self):
        self._check(load().fmm_downward(self._h), "fmm_downward")

    def evaluate(self, graph=False):
        """one solve; graph=True replays the solve captured as a HIP graph
        (captured on the first such call after the bodies change)"""
        if graph:
            self._check(load().fmm_evaluate_graph(self._h), "fmm_evaluate_graph")
        else:
            self._check(load().fmm_evaluate(self._h), "fmm_evaluate")

    def synchronize(self):
        self._check(load().fmm_synchronize(self._h), "fmm_synchronize")

    def targets(self, input_order=True):
        trg = np.zeros((self.n, 4))
        self._check(load().fmm_get_targets(self._h, _ptr(trg), 1 if input_order else 0),
                    "fmm_get_targets")
        return trg

    def direct(self, xt):
        """Traversal::direct at the points xt against all bodies (GPU)."""
        xt = np.ascontiguousarray(xt, dtype=np.float64).reshape(-1, 3)
        trg = np.zeros((len(xt), 4))
        self._check(load().fmm_direct(self._h, len(xt), _ptr(xt), _ptr(trg)), "fmm_direct")
        return trg

    def dipole_correction(self):
        """UpDownPass::getDipole + dipoleCorrection on the device targets."""
        self._check(load().fmm_dipole_correction(self._h), "fmm_dipole_correction")

    def md_evaluate(self, x, q, p=None, f=None):
        """FMM_FMM (wrappers/md_distributed.cxx:190-264) on flat arrays: returns
        (p, f) with the periodic potential and acceleration added to p, f."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64)
        n = len(q)
        if len(x) != n:
            raise ValueError("x and q lengths differ")
        p = np.zeros(n) if p is None else np.array(p, dtype=np.float64)
        f = np.zeros((n, 3)) if f is None else np.array(f, dtype=np.float64).reshape(n, 3)
        self._check(load().fmm_md_evaluate(self._h, n, _ptr(x), _ptr(q), _ptr(p), _ptr(f)),
                    "fmm_md_evaluate")
        self.n = n
        return p, f

    def ewald(self, x, q, ksize, alpha, sigma, cutoff, p=None, f=None):
        """FMM_Ewald (wrappers/md_distributed.cxx:310-360) on the device: returns
        (p, f) with the Ewald potential and gradient added, period = cycle."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64)
        n = len(q)
        if len(x) != n:
            raise ValueError("x and q lengths differ")
        p = np.zeros(n) if p is None else np.array(p, dtype=np.float64)
        f = np.zeros((n, 3)) if f is None else np.array(f, dtype=np.float64).reshape(n, 3)
        self._check(load().fmm_ewald(self._h, n, _ptr(x), _ptr(q), ksize, alpha, sigma, cutoff, _ptr(p),
                                     _ptr(f)), "fmm_ewald")
        return p, f

    def ewald_noalloc(self, n, x_dev, q_dev, ksize, alpha, sigma, cutoff, p_dev, f_dev):
        """fmm_ewald_noalloc: the Ewald sum on device arrays given by address
        (x n x 3, q n, p n, f n x 3, float64, on this context's device); adds
        into p and f on the context's stream without waiting and allocates
        nothing (ksize <= 32)."""
        self._check(load().fmm_ewald_noalloc(self._h, int(n), x_dev, q_dev, int(ksize), alpha, sigma, cutoff,
                                             p_dev, f_dev), "fmm_ewald_noalloc")

    def stats(self):
        s = Stats()
        self._check(load().fmm_get_stats(self._h, ctypes.byref(s)), "fmm_get_stats")
        return s.as_dict()

    def expansions(self):
        nc = self.stats()["ncell"]
        M = np.zeros((nc, self.nterm), np.complex128)
        Lc = np.zeros((nc, self.nterm), np.complex128)
        self._check(load().fmm_get_expansions(self._h, _ptr(M.view(np.float64)),
                                              _ptr(Lc.view(np.float64))), "fmm_get_expansions")
        return M, Lc

    def tree(self):
        s = self.stats()
        nc, n = s["ncell"], s["n"]
        a = {k: np.zeros(nc, np.int32) for k in ("iparent", "ichild", "nchild", "ibody", "nbody")}
        perm = np.zeros(n, np.int32)
        X = np.zeros((nc, 3))
        R = np.zeros(nc)
        self._check(load().fmm_get_tree(self._h, _ptr(perm, _ip), _ptr(a["iparent"], _ip),
                                        _ptr(a["ichild"], _ip), _ptr(a["nchild"], _ip),
                                        _ptr(a["ibody"], _ip), _ptr(a["nbody"], _ip), _ptr(X),
                                        _ptr(R)), "fmm_get_tree")
        a.update(perm=perm, X=X, R=R, depth=s["depth"])
        return a

    def lists(self):
        s = self.stats()
        nc = s["ncell"]
        m2l_ptr = np.zeros(nc + 1, np.int32)
        p2p_ptr = np.zeros(nc + 1, np.int32)
        m2l_src = np.zeros(max(s["m2l_pairs"], 1), np.int32)
        p2p_src = np.zeros(max(s["p2p_pairs"], 1), np.int32)
        self._check(load().fmm_get_lists(self._h, _ptr(m2l_ptr, _ip), _ptr(m2l_src, _ip),
                                         _ptr(p2p_ptr, _ip), _ptr(p2p_src, _ip)), "fmm_get_lists")
        return m2l_ptr, m2l_src[:s["m2l_pairs"]], p2p_ptr, p2p_src[:s["p2p_pairs"]]

    def list_counts(self):
        """Traversal::countList per cell: (ICELL Morton keys, numP2P, numM2L)."""
        nc = self.stats()["ncell"]
        key = np.zeros(max(nc, 1), np.uint64)
        p2p, m2l = np.zeros(max(nc, 1), np.int32), np.zeros(max(nc, 1), np.int32)
        self._check(load().fmm_get_list_counts(self._h, key.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                               _ptr(p2p, _ip), _ptr(m2l, _ip)), "fmm_get_list_counts")
        return key[:nc], p2p[:nc], m2l[:nc]


    # ---- Kernel operators (kernels/laplace.h:150-388) on caller-held cells:
    # expansions are complex arrays of NTERM (reference normalisation); each
    # call returns the input accumulator plus the operator's result
    def _c3(self, v):
        return np.ascontiguousarray(v, dtype=np.float64).reshape(-1)

    def _cx(self, E, shape):
        E = np.zeros(shape, np.complex128) if E is None else np.array(E, dtype=np.complex128, order="C")
        return E

    def kernel_p2p(self, xi, xj, qj, xper=None, trg=None):
        xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
        xj = np.ascontiguousarray(xj, dtype=np.float64).reshape(-1, 3)
        qj = np.ascontiguousarray(qj, dtype=np.float64).reshape(-1)
        trg = np.zeros((len(xi), 4)) if trg is None else np.array(trg, dtype=np.float64, order="C")
        xp = None if xper is None else self._c3(xper)
        self._check(load().fmm_kernel_p2p(self._h, len(xi), _ptr(xi), _ptr(trg), len(xj), _ptr(xj), _ptr(qj),
                                          _ptr(xp) if xp is not None else None), "fmm_kernel_p2p")
        return trg

    def kernel_p2m(self, Xc, x, q, M=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        M = self._cx(M, self.nterm)
        self._check(load().fmm_kernel_p2m(self._h, _ptr(self._c3(Xc)), len(x), _ptr(x), _ptr(q),
                                          _ptr(M.view(np.float64))), "fmm_kernel_p2m")
        return M

    def kernel_m2m(self, Xp, Xc, Mc, Mp=None):
        Xc = np.ascontiguousarray(Xc, dtype=np.float64).reshape(-1, 3)
        Mc = np.ascontiguousarray(Mc, dtype=np.complex128).reshape(len(Xc), self.nterm)
        Mp = self._cx(Mp, self.nterm)
        self._check(load().fmm_kernel_m2m(self._h, _ptr(self._c3(Xp)), len(Xc), _ptr(Xc),
                                          _ptr(Mc.view(np.float64)), _ptr(Mp.view(np.float64))), "fmm_kernel_m2m")
        return Mp

    def kernel_m2l(self, Xi, Xj, Mj, Li=None, xper=None):
        """Batched: Xi, Xj (k, 3), Mj (k, NTERM) -> Li (k, NTERM)."""
        Xi = np.ascontiguousarray(Xi, dtype=np.float64).reshape(-1, 3)
        Xj = np.ascontiguousarray(Xj, dtype=np.float64).reshape(-1, 3)
        k = len(Xi)
        Mj = np.ascontiguousarray(Mj, dtype=np.complex128).reshape(k, self.nterm)
        Li = self._cx(Li, (k, self.nterm))
        xp = None if xper is None else self._c3(xper)
        self._check(load().fmm_kernel_m2l(self._h, k, _ptr(Xi), _ptr(Xj), _ptr(xp) if xp is not None else None,
                                          _ptr(Mj.view(np.float64)), _ptr(Li.view(np.float64))), "fmm_kernel_m2l")
        return Li

    def kernel_l2l(self, Xc, Xp, Lp, Lc=None):
        Lp = np.ascontiguousarray(Lp, dtype=np.complex128).reshape(self.nterm)
        Lc = self._cx(Lc, self.nterm)
        self._check(load().fmm_kernel_l2l(self._h, _ptr(self._c3(Xc)), _ptr(self._c3(Xp)),
                                          _ptr(Lp.view(np.float64)), _ptr(Lc.view(np.float64))), "fmm_kernel_l2l")
        return Lc

    def kernel_l2p(self, Xc, L, x, trg=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        L = np.ascontiguousarray(L, dtype=np.complex128).reshape(self.nterm)
        trg = np.zeros((len(x), 4)) if trg is None else np.array(trg, dtype=np.float64, order="C")
        self._check(load().fmm_kernel_l2p(self._h, _ptr(self._c3(Xc)), _ptr(L.view(np.float64)), len(x), _ptr(x),
                                          _ptr(trg)), "fmm_kernel_l2p")
        return trg


class Plan:
    """Tree + interaction lists (fmm_plan_*), built on the host, or with the
    lists built on the device of an FMM context (lists_on=)."""

    def __init__(self, x, ncrit=64, theta=0.4, nspawn=5000, tree=None, lists_on=None):
        """Build the tree of x, or, with tree= (dict of iparent, ichild, nchild,
        ibody, nbody, X, R; x already in tree order), adopt that tree.
        lists_on=FMM: the lists from that context's device (fmm_plan_build_device)."""
        L = load()
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self.x = x
        self._h = ctypes.c_void_p()
        if lists_on is not None:
            rc = L.fmm_plan_build_device(lists_on._h, ctypes.byref(self._h), len(x), _ptr(x), ncrit, theta,
                                         nspawn)
            lists_on._check(rc, "fmm_plan_build_device")
            return
        if tree is None:
            rc = L.fmm_plan_build(ctypes.byref(self._h), len(x), _ptr(x), ncrit, theta, nspawn)
            if rc != 0:
                raise FMMError("fmm_plan_build failed (rc=%d)" % rc)
            return
        a = {k: np.ascontiguousarray(tree[k], dtype=np.int32)
             for k in ("iparent", "ichild", "nchild", "ibody", "nbody")}
        X = np.ascontiguousarray(tree["X"], dtype=np.float64).reshape(-1, 3)
        R = np.ascontiguousarray(tree["R"], dtype=np.float64).reshape(-1)
        why = ctypes.create_string_buffer(256)
        rc = L.fmm_plan_from_tree(ctypes.byref(self._h), len(x), _ptr(x), len(R), _ptr(a["iparent"], _ip),
                                  _ptr(a["ichild"], _ip), _ptr(a["nchild"], _ip), _ptr(a["ibody"], _ip),
                                  _ptr(a["nbody"], _ip), _ptr(X), _ptr(R), theta, nspawn, why, 256)
        if rc != 0:
            raise FMMError("fmm_plan_from_tree failed (rc=%d): %s" % (rc, why.value.decode()))

    def __del__(self):
        try:
            if self._h:
                load().fmm_plan_destroy(self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass

    @classmethod
    def tree_in_bounds(cls, x, ncrit, xmin, xmax):
        """BuildTree::buildTree(bodies, buffer, bounds): the octree of x in the
        given bounds, no lists (fmm_plan_build_tree)."""
        L = load()
        self = cls.__new__(cls)
        self.x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        self._h = ctypes.c_void_p()
        lo = np.ascontiguousarray(xmin, dtype=np.float64)
        hi = np.ascontiguousarray(xmax, dtype=np.float64)
        rc = L.fmm_plan_build_tree(ctypes.byref(self._h), len(self.x), _ptr(self.x), ncrit, _ptr(lo), _ptr(hi))
        if rc != 0:
            raise FMMError("fmm_plan_build_tree failed (rc=%d)" % rc)
        return self

    @classmethod
    def pair(cls, targets, sources, theta=0.4, nspawn=5000):
        """Traversal::traverse(icells, jcells): the forest of a target plan and
        a source plan with the lists of the pair traversal (fmm_plan_pair);
        bodies are the targets' x then the sources' x."""
        L = load()
        self = cls.__new__(cls)
        self.x = np.concatenate([targets.x, sources.x])
        self._h = ctypes.c_void_p()
        rc = L.fmm_plan_pair(ctypes.byref(self._h), targets._h, sources._h, theta, nspawn)
        if rc != 0:
            raise FMMError("fmm_plan_pair failed (rc=%d)" % rc)
        return self

    @classmethod
    def merge(cls, plans):
        """Forest of independent systems (fmm_plan_merge); bodies are the
        plans' x concatenated in order."""
        L = load()
        self = cls.__new__(cls)
        self.x = np.concatenate([p.x for p in plans]) if plans else np.zeros((0, 3))
        self._h = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(plans))(*[p._h.value for p in plans])
        rc = L.fmm_plan_merge(ctypes.byref(self._h), len(plans), arr)
        if rc != 0:
            raise FMMError("fmm_plan_merge failed (rc=%d)" % rc)
        return self

    def prepare(self, P=10, precision=FMM_FP64, m2l_mode=M2L_AUTO):
        """Build the M2L groups (and target tiles) a context with these
        parameters needs (fmm_plan_prepare)."""
        L = load()
        p = Params()
        L.fmm_default_params(ctypes.byref(p))
        p.P, p.precision, p.m2l_mode = P, precision, m2l_mode
        rc = L.fmm_plan_prepare(self._h, ctypes.byref(p))
        if rc != 0:
            raise FMMError("fmm_plan_prepare failed (rc=%d)" % rc)

    def build_groups(self, precision=FMM_FP64):
        rc = load().fmm_plan_build_groups(self._h, precision)
        if rc != 0:
            raise FMMError("fmm_plan_build_groups failed (rc=%d)" % rc)

    def stats(self):
        s = Stats()
        load().fmm_plan_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def tree(self):
        s = self.stats()
        nc, n = s["ncell"], s["n"]
        a = {k: np.zeros(nc, np.int32) for k in ("iparent", "ichild", "nchild", "ibody", "nbody")}
        perm = np.zeros(n, np.int32)
        X = np.zeros((nc, 3))
        R = np.zeros(nc)
        load().fmm_plan_tree(self._h, _ptr(perm, _ip), _ptr(a["iparent"], _ip),
                             _ptr(a["ichild"], _ip), _ptr(a["nchild"], _ip), _ptr(a["ibody"], _ip),
                             _ptr(a["nbody"], _ip), _ptr(X), _ptr(R))
        a.update(perm=perm, X=X, R=R, depth=s["depth"])
        return a

    def lists(self):
        s = self.stats()
        nc = s["ncell"]
        m2l_ptr = np.zeros(nc + 1, np.int32)
        p2p_ptr = np.zeros(nc + 1, np.int32)
        m2l_src = np.zeros(max(s["m2l_pairs"], 1), np.int32)
        p2p_src = np.zeros(max(s["p2p_pairs"], 1), np.int32)
        rc = load().fmm_plan_lists(self._h, _ptr(m2l_ptr, _ip), _ptr(m2l_src, _ip), _ptr(p2p_ptr, _ip),
                                   _ptr(p2p_src, _ip))
        if rc != 0:
            raise FMMError("fmm_plan_lists failed (rc=%d)" % rc)
        return m2l_ptr, m2l_src[:s["m2l_pairs"]], p2p_ptr, p2p_src[:s["p2p_pairs"]]

    def list_counts(self):
        """Traversal::countList per cell: (ICELL Morton keys, numP2P, numM2L)."""
        nc = self.stats()["ncell"]
        key = np.zeros(max(nc, 1), np.uint64)
        p2p, m2l = np.zeros(max(nc, 1), np.int32), np.zeros(max(nc, 1), np.int32)
        load().fmm_plan_list_counts(self._h, key.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                    _ptr(p2p, _ip), _ptr(m2l, _ip))
        return key[:nc], p2p[:nc], m2l[:nc]


class Server:
    """Request loop (fmm_server_*): submit systems as they arrive, wait for
    each result.  Keyword arguments of FMM give the solve parameters;
    host_threads, slots, device_budget, batch_bodies, batch_requests the loop."""

    def __init__(self, P=8, theta=0.4, ncrit=64, nspawn=5000, eps2=0.0, device=0,
                 precision=FMM_FP64, m2l_mode=M2L_AUTO, **cfg):
        L = load()
        p = Params()
        L.fmm_default_params(ctypes.byref(p))
        p.P, p.theta, p.ncrit, p.nspawn, p.eps2 = P, theta, ncrit, nspawn, eps2
        p.device, p.precision, p.m2l_mode = device, precision, m2l_mode
        c = ServerConfig()
        L.fmm_server_default_config(ctypes.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self._h = ctypes.c_void_p()
        self._n = {}
        rc = L.fmm_server_create(ctypes.byref(self._h), ctypes.byref(p), ctypes.byref(c))
        if rc != 0:
            self._h = ctypes.c_void_p()
            raise FMMError("fmm_server_create failed (rc=%d)" % rc)

    def submit(self, x, q):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        rid = ctypes.c_int64()
        rc = load().fmm_server_submit(self._h, len(x), _ptr(x), _ptr(q), ctypes.byref(rid))
        if rc != 0:
            raise FMMError("fmm_server_submit failed (rc=%d)" % rc)
        self._n[rid.value] = len(x)
        return rid.value

    def wait(self, rid):
        """(targets (n,4) in input order, request info dict)"""
        trg = np.zeros((self._n.pop(rid), 4))
        info = RequestInfo()
        rc = load().fmm_server_wait(self._h, rid, _ptr(trg), ctypes.byref(info))
        if rc != 0:
            raise FMMError("fmm_server_wait failed (rc=%d)" % rc)
        return trg, info.as_dict()

    def stats(self):
        s = ServerStats()
        load().fmm_server_get_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def close(self):
        if self._h:
            load().fmm_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
