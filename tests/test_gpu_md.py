"""The MD-style periodic entry point (fmm_md_evaluate: FMM_FMM of
wrappers/md_distributed.cxx:190-264 on one process) and the dipole correction
(include/up_down_pass.h:77-99), on the GPU, against the oracle; and the
periodic FMM against the Ewald sum (include/ewald.h) the way
wrappers/test_md_distributed.cxx:78-108 validates it (energy sum q phi and
acceleration, relative L2).

Bars: as for the periodic FMM (acceleration and de-meaned potential 1e-10,
the potential's constant 1e-7 of its rms); the FMM-vs-Ewald errors of the GPU
equal the oracle's to 1e-6 of their value.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

L = 2 * np.pi


def _fmm():
    import exafmm_amd
    return exafmm_amd


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _system(n, seed, spread=0.5):
    """zero-mean charges; positions in [-spread L, spread L]^3 (beyond the box when spread > 1/2)"""
    rng = np.random.default_rng(seed)
    x = (rng.random((n, 3)) * 2 - 1) * spread * L
    q = rng.random(n) - 0.5
    return x, q - q.mean()


def _oracle_md(x, q, P, theta, images, ncrit=32, nspawn=1000):
    xw, _ = O.wrap(x, L)
    t = O.Tree(xw, ncrit)
    ref = t.fmm_periodic(O.Kernel(P), xw[t.perm], q[t.perm], theta, nspawn, images, L)[np.argsort(t.perm)]
    return O.dipole_correction(xw, q, L, np.ascontiguousarray(ref))


def _check(p, f, ref):
    r = ref[:, 0]
    assert _rel(p - p.mean(), r - r.mean()) < 1e-10
    assert _rel(f, ref[:, 1:]) < 1e-10
    assert abs(p.mean() - r.mean()) < 1e-7 * np.sqrt(np.mean(r * r))


@pytest.mark.parametrize("P,theta,images", [(10, 0.4, 3), (8, 0.6, 2), (6, 0.4, 1)])
def test_md_evaluate_matches_oracle(P, theta, images):
    print(f"md P={P} theta={theta} images={images}", flush=True)
    x, q = _system(3000, 1, spread=0.7)  # about 2/3 of the bodies start outside the box
    rng = np.random.default_rng(2)
    p0, f0 = rng.random(3000), rng.random((3000, 3))
    with _fmm().FMM(P=P, theta=theta, ncrit=32, nspawn=1000, images=images, cycle=L) as f:
        p, fa = f.md_evaluate(x, q, p0, f0)
    ref = _oracle_md(x, q, P, theta, images)
    _check(p - p0, fa - f0, ref)


def test_dipole_correction_closed_form():
    x, q = _system(5000, 3)
    q = q + 0.05 * np.sign(x[:, 0])  # a clear net dipole along x
    with _fmm().FMM(P=6, images=1, cycle=L) as f:
        f.set_bodies(x, q)
        f.evaluate()
        before = f.targets()
        f.dipole_correction()
        after = f.targets()
    D = (x * q[:, None]).sum(0)
    coef = 4 * np.pi / (3 * L ** 3)
    assert np.allclose(after[:, 1:] - before[:, 1:], -coef * D, rtol=1e-12, atol=1e-15)
    assert np.allclose(after[:, 0] - before[:, 0], -coef * (D @ D) / len(q) / q, rtol=1e-11, atol=0)


@pytest.mark.parametrize("P,theta,images", [(10, 0.4, 3), (12, 0.6, 4)])
def test_md_against_ewald(P, theta, images):
    """wrappers/test_md_distributed.cxx's check (ksize 11, alpha 10/L,
    sigma 1/(4 pi), cutoff L/2): the GPU's errors against Ewald are the
    oracle's.  They are large for the reference's algorithm itself: at
    theta = 0.4 its image set leaves the shell at distance 2 out (the near
    part is the 27 boxes, the first supercell level starts at shell 3); at
    theta = 0.6 shell 2 is reached by M2L from whole-root copies at a
    convergence ratio of about 0.87.  The energy error is relative to an
    Ewald energy that is small beside the dipole term for random charges."""
    n = 2000
    x, q = _system(n, 4)
    ew = O.ewald(x, q, 11, 10 / L, 0.25 / np.pi, L / 2, L)
    with _fmm().FMM(P=P, theta=theta, ncrit=32, nspawn=1000, images=images, cycle=L) as f:
        p, fa = f.md_evaluate(x, q)
    ref = _oracle_md(x, q, P, theta, images)
    e_pot = lambda pp: abs(q @ pp - q @ ew[:, 0]) / abs(q @ ew[:, 0])
    e_acc = lambda ff: _rel(ff, ew[:, 1:])
    print(f"P={P} theta={theta} images={images}: energy {e_pot(p):.3e} acc {e_acc(fa):.3e}", flush=True)
    assert abs(e_pot(p) - e_pot(ref[:, 0])) < 1e-6 * e_pot(ref[:, 0])
    assert abs(e_acc(fa) - e_acc(ref[:, 1:])) < 1e-6 * e_acc(ref[:, 1:])


def test_md_evaluate_empty_and_bad_arguments():
    with _fmm().FMM(P=6, images=1, cycle=L) as f:
        p, fa = f.md_evaluate(np.zeros((0, 3)), np.zeros(0))
        assert p.shape == (0,) and fa.shape == (0, 3)
        with pytest.raises(ValueError):
            f.md_evaluate(np.zeros((3, 3)), np.zeros(2))


def test_device_ewald_matches_oracle():
    """fmm_ewald (FMM_Ewald on the device) equals the oracle's Ewald sum, with
    bodies outside the box wrapped and the caller's p, f accumulated."""
    x, q = _system(1500, 6, spread=0.6)
    rng = np.random.default_rng(7)
    p0, f0 = rng.random(1500), rng.random((1500, 3))
    with _fmm().FMM(P=6, images=1, cycle=L) as f:
        p, fa = f.ewald(x, q, 11, 10 / L, 0.25 / np.pi, L / 2, p0, f0)
    xw, _ = O.wrap(x, L)
    ref = O.ewald(xw, q, 11, 10 / L, 0.25 / np.pi, L / 2, L)
    assert _rel(p - p0, ref[:, 0]) < 1e-12
    assert _rel(fa - f0, ref[:, 1:]) < 1e-12


def test_device_ewald_rock_salt():
    """The rock-salt Madelung constant through the device Ewald sum."""
    g = np.arange(4)
    i, j, k = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([i, j, k], -1).reshape(-1, 3).astype(float) - 1.5
    q = np.where((i + j + k).reshape(-1) % 2 == 0, 1.0, -1.0)
    with _fmm().FMM(P=6, images=1, cycle=4.0) as f:
        p, fa = f.ewald(x, q, 24, 3.0, 0.25 / np.pi, 2.0)
    assert np.max(np.abs(q * p + 1.747564594633182)) < 1e-10
    assert np.max(np.abs(fa)) < 1e-10
