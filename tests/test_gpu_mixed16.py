"""Mixed precision with the M2L on split-FP16 operands (FMM_MIXED16: hi + lo
halves of every power-of-two-scaled operator row and cell multipole,
v_mfma_f32_16x16x32_f16 as hi x hi + hi x lo + lo x hi, FP32 accumulation,
FP64 across lists) against the FP64 path, the FP32-operand mixed path and the
FP64 direct sum.

Tolerances as for FMM_MIXED: rel-L2 distance to the FP64 FMM below 1e-6 for
the potential and 1e-5 for the acceleration (the FP32 P2P's position
rounding), and an error against the direct sum within 1e-6 / 5e-6 of the FP64
FMM's own.  A CPU emulation of the split operands on the operators of P = 12
gave a per-interaction potential error of 1.7e-7 (FP32 operands: 4e-8).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - b) / np.linalg.norm(b))


def _solve(F, x, q, P, precision, **kw):
    f = F.FMM(P=P, precision=precision, **kw)
    f.set_bodies(x, q)
    f.evaluate()
    out = f.targets(input_order=True), f.stats()
    f.close()
    return out


@pytest.mark.parametrize("P,dist,n", [(4, "cube1", 20000), (8, "cube1", 40000), (10, "plummer1", 40000),
                                      (11, "cube1", 30000), (12, "cube1", 40000), (12, "plummer1", 30000)])
def test_mixed16_matches_fp64(P, dist, n):
    import exafmm_amd as F
    print("mixed16 P=%d %s n=%d" % (P, dist, n), flush=True)
    x, q = F.make_bodies(n, dist, 1)
    t64, _ = _solve(F, x, q, P, F.FMM_FP64)
    t32, _ = _solve(F, x, q, P, F.FMM_MIXED)
    t16, s16 = _solve(F, x, q, P, F.FMM_MIXED16)
    assert s16["m2l_path"] == F.M2L_TILES and s16["m2l_groups"] > 0
    dp, da = _rel(t16[:, 0], t64[:, 0]), _rel(t16[:, 1:], t64[:, 1:])
    d32 = _rel(t32[:, 0], t64[:, 0])
    idx = np.arange(1000) * (n // 1000)
    d = O.direct(O.Kernel(P), x[idx], x, q)
    e64, e16 = O.rel_l2(d, t64[idx]), O.rel_l2(d, t16[idx])
    print("P=%d %s: mixed16-fp64 %.2e %.2e (mixed-fp64 pot %.2e); vs direct fp64 %.2e %.2e mixed16 %.2e %.2e"
          % (P, dist, dp, da, d32, e64[0], e64[1], e16[0], e16[1]))
    # worst single target, against the FP32-operand path's worst: the split
    # operands must not leave outliers (a scale per cell instead of per
    # (cell, k-block) flushed high-degree components of some interactions)
    rms = float(np.sqrt(np.mean(t64[:, 0] ** 2)))
    w16 = float(np.max(np.abs(t16[:, 0] - t64[:, 0]))) / rms
    w32 = float(np.max(np.abs(t32[:, 0] - t64[:, 0]))) / rms
    print("   worst target potential / rms: mixed16 %.2e, mixed %.2e" % (w16, w32))
    assert dp < 1e-6 and da < 1e-5
    assert e16[0] < e64[0] + 1e-6 and e16[1] < e64[1] + 5e-6
    assert w16 < 3 * w32 + 1e-6


def test_mixed16_periodic_and_refusals():
    import exafmm_amd as F
    L = 2 * np.pi
    x, q = F.make_bodies(3000, "cube1", 2)
    x = x * (L / 2 * 0.999)
    kw = dict(P=8, images=2, cycle=L)
    t64, _ = _solve(F, x, q, precision=F.FMM_FP64, **kw)
    t16, _ = _solve(F, x, q, precision=F.FMM_MIXED16, **kw)
    assert _rel(t16[:, 1:], t64[:, 1:]) < 1e-5
    for bad in (dict(P=13), dict(P=8, m2l_mode=F.M2L_GROUPED), dict(P=8, m2l_mode=F.M2L_TARGET)):
        with pytest.raises(RuntimeError):
            F.FMM(precision=F.FMM_MIXED16, **bad)
