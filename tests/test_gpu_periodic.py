"""GPU parity of the periodic path (images > 0) against the oracle's periodic
FMM (the 27 shifted dual traversals plus traversePeriodic's supercell levels,
include/traversal.h:233-347) and of the periodic direct sum against brute
force over the shifts of Traversal::direct (include/traversal.h:355-377).

Bar: 1e-10 relative (L2) between the GPU FP64 FMM and the oracle on the same
tree and lists, for the acceleration and for the potential less its mean.  The
far images enter the GPU through one pre-composed operator per system instead
of the reference's level-by-level M2M/M2L, so the two sides agree to rounding
only.  The potential's constant part is ill-conditioned in FP64: the root's
net charge is zero only to rounding (|sum q| ~ 1e-15 for these systems) and
the image sum of 1/r multiplies it by ~(5^(images-1) * 2 pi)^2, so at
images >= 3 the constant moves by ~1e-9 relative under any change of
summation order (a long-double build of the oracle puts the FP64 oracle's own
constant 5e-9 off at P = 12, images = 4, while the de-meaned potentials agree
to 2e-15).  That constant is bounded separately (OFFSET_TOL).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FMM_TOL = 1e-10
OFFSET_TOL = 1e-7  # constant potential offset / rms potential


def _check_periodic(trg, ref):
    """acceleration and de-meaned potential to FMM_TOL; the constant to OFFSET_TOL"""
    p, r = trg[:, 0], ref[:, 0]
    assert _rel(p - p.mean(), r - r.mean()) < FMM_TOL
    assert _rel(trg[:, 1:], ref[:, 1:]) < FMM_TOL
    assert abs(p.mean() - r.mean()) < OFFSET_TOL * np.sqrt(np.mean(r * r))


def _fmm():
    import exafmm_amd
    return exafmm_amd


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _periodic_pair(n, P, images, theta, seed=0, cyc=2 * np.pi, ncrit=64, **kw):
    x, q = O.data_cube(n, seed), O.data_source(n, seed)
    f = _fmm().FMM(P=P, theta=theta, ncrit=ncrit, images=images, cycle=cyc, **kw)
    f.set_bodies(x, q)
    f.evaluate()
    t = O.Tree(x, ncrit)
    xs, qs = x[t.perm], q[t.perm]
    ref = t.fmm_periodic(O.Kernel(P), xs, qs, theta, 5000, images, cyc)
    return f, ref, x, q


@pytest.mark.parametrize("P,images,theta,n", [(6, 1, 0.4, 3000), (10, 1, 0.4, 3000), (10, 2, 0.4, 3000),
                                              (8, 3, 0.6, 4000), (10, 3, 0.35, 3000), (12, 4, 0.4, 2000)])
def test_periodic_fmm_matches_oracle(P, images, theta, n):
    print(f"periodic P={P} images={images} theta={theta} n={n}", flush=True)
    f, ref, _, _ = _periodic_pair(n, P, images, theta)
    trg = f.targets(input_order=False)
    f.close()
    _check_periodic(trg, ref)


def test_periodic_rectangular_cycle():
    """cycle differs per axis: the image shifts scale each coordinate on its own."""
    cyc = (2 * np.pi, 2.5 * np.pi, 3 * np.pi)
    f, ref, _, _ = _periodic_pair(2500, 8, 2, 0.4, seed=3, cyc=cyc)
    trg = f.targets(input_order=False)
    f.close()
    _check_periodic(trg, ref)


@pytest.mark.parametrize("mode", [1, 2])
def test_periodic_m2l_layouts_agree(mode):
    """Target tiles and pair chunks give the same periodic result."""
    x, q = O.data_cube(6000, 7), O.data_source(6000, 7)
    out = []
    for m in (0, mode):
        f = _fmm().FMM(P=8, theta=0.4, images=2, cycle=2 * np.pi, m2l_mode=m)
        f.set_bodies(x, q)
        f.evaluate()
        out.append(f.targets())
        f.close()
    assert _rel(out[1], out[0]) < 1e-13


def test_periodic_mixed_precision_close_to_fp64():
    x, q = O.data_cube(20000, 11), O.data_source(20000, 11)
    out = []
    for prec in (_fmm().FMM_FP64, _fmm().FMM_MIXED):
        f = _fmm().FMM(P=8, theta=0.4, images=3, cycle=2 * np.pi, precision=prec)
        f.set_bodies(x, q)
        f.evaluate()
        out.append(f.targets())
        f.close()
    assert _rel(out[1][:, 0], out[0][:, 0]) < 1e-6
    assert _rel(out[1][:, 1:], out[0][:, 1:]) < 1e-5


@pytest.mark.parametrize("images,theta", [(1, 0.4), (2, 0.4), (2, 0.6)])
def test_periodic_direct_matches_shift_sum(images, theta):
    """fmm_direct with images: every shift (ix, iy, iz) * cycle with
    |i| <= sum_(l < images) neighbor^l, neighbor = int(2 (.5/theta + 1) + 1)."""
    n = 200
    x, q = O.data_cube(n, 2), O.data_source(n, 2)
    cyc = 2 * np.pi
    f = _fmm().FMM(P=6, theta=theta, images=images, cycle=cyc)
    f.set_bodies(x, q)
    got = f.direct(x)
    f.close()
    neighbor = int(2 * (0.5 / theta + 1) + 1)
    prange = sum(neighbor ** l for l in range(images))
    k = O.Kernel(6)
    ref = np.zeros((n, 4))
    rng = range(-prange, prange + 1)
    for ix in rng:
        for iy in rng:
            for iz in rng:
                ref += O.direct(k, x, x + np.array([ix, iy, iz]) * cyc, q)
    assert _rel(got[:, 0], ref[:, 0]) < 1e-12
    assert _rel(got[:, 1:], ref[:, 1:]) < 1e-12


def test_periodic_fmm_tracks_its_direct_sum():
    """images = 1 truncates at the 27 boxes both in the FMM and in the direct
    sum: the GPU's FMM-vs-direct error is the oracle's (the expansion error of
    P = 12, theta = 0.4: 1.24e-6 potential, 5.6e-7 acceleration)."""
    n, cyc = 2000, 2 * np.pi
    x, q = O.data_cube(n, 5), O.data_source(n, 5)
    f = _fmm().FMM(P=12, theta=0.4, images=1, cycle=cyc)
    f.set_bodies(x, q)
    f.evaluate()
    trg = f.targets()
    idx = np.arange(0, n, 7)
    d = f.direct(x[idx])
    f.close()
    t = O.Tree(x, 64)
    k = O.Kernel(12)
    ref = t.fmm_periodic(k, x[t.perm], q[t.perm], 0.4, 5000, 1, cyc)[np.argsort(t.perm)]
    dref = np.zeros((len(idx), 4))
    for a in (-1, 0, 1):
        for b in (-1, 0, 1):
            for c in (-1, 0, 1):
                dref += O.direct(k, x[idx], x + np.array([a, b, c]) * cyc, q)
    for sl in (slice(0, 1), slice(1, 4)):
        e_gpu, e_ref = _rel(trg[idx, sl], d[:, sl]), _rel(ref[idx, sl], dref[:, sl])
        assert e_ref < 2e-6 and abs(e_gpu - e_ref) < 1e-6 * e_ref, (e_gpu, e_ref)
