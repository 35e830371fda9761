"""fmm_ewald_noalloc (the allocation-free Ewald entry point) against fmm_ewald.

Both run the same kernels in the same summation order, so on the same bodies
(some outside the box: both wrap each position once) the results are equal
bit for bit when the caller's outputs start at zero.  Two calls in a row, the
second with fewer bodies and a smaller ksize, check that nothing of the first
call's scratch contents leaks into the second; a ksize beyond the scratch
limit is refused.  The caller's device arrays are allocated through the HIP
runtime the library itself uses (found among the process's mapped objects),
so no second runtime is loaded.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 2 * np.pi


def _system(n, seed):
    rng = np.random.default_rng(seed)
    x = (rng.random((n, 3)) * 2 - 1) * 0.6 * L  # about a sixth of each coordinate beyond the box
    q = rng.random(n) - 0.5
    return x, q - q.mean()


def _hip():
    with open("/proc/self/maps") as fh:
        paths = {l.split()[-1] for l in fh if "libamdhip64" in l}
    assert paths, "the library's HIP runtime is not mapped"
    h = ctypes.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return h


class _Dev:
    """a device copy of a float64 array (hipMemcpy kinds: 1 host->device, 2 device->host)"""

    def __init__(self, hip, a):
        self.hip, self.a = hip, np.ascontiguousarray(a, dtype=np.float64)
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_ewald_noalloc_matches_ewald():
    import exafmm_amd as F
    with F.FMM(P=6, images=1, cycle=L) as f:
        hip = _hip()
        for n, ksize, seed in ((3000, 12, 1), (1000, 8, 2)):
            print("ewald_noalloc n=%d ksize=%d" % (n, ksize), flush=True)
            x, q = _system(n, seed)
            alpha = 10.0 / L
            pe, fe = f.ewald(x, q, ksize, alpha, 0.25 / np.pi, L / 2)
            xd, qd = _Dev(hip, x), _Dev(hip, q)
            pd, fd = _Dev(hip, np.zeros(n)), _Dev(hip, np.zeros((n, 3)))
            f.ewald_noalloc(n, xd.p, qd.p, ksize, alpha, 0.25 / np.pi, L / 2, pd.p, fd.p)
            f.synchronize()
            p, g = pd.get(), fd.get()
            assert np.array_equal(p, pe), np.max(np.abs(p - pe))
            assert np.array_equal(g, fe), np.max(np.abs(g - fe))
            if n == 1000:  # beyond the scratch array's limit: refused before any launch
                with pytest.raises(RuntimeError, match="scratch limit"):
                    f.ewald_noalloc(1, xd.p, qd.p, 33, alpha, 0.25 / np.pi, L / 2, pd.p, fd.p)
            for d in (xd, qd, pd, fd):
                d.free()
