"""fmm_evaluate_graph: the solve captured once as a HIP graph and replayed.

Every replay is checked against the eager fmm_evaluate of the same system,
with eager work (solves, direct sums, charge uploads) and synchronisations
between the replays.  The graph reads the charges in place, so after
fmm_set_charges a replay solves the new charges; after fmm_set_bodies it is
captured again.  Bars: the eager and the captured solve run the same
kernels; split M2L tiles add their parts with atomics, so results agree to
rounding (1e-13 relative), and the first system also against the oracle
(1e-10).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-13


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _close(a, b):
    return _rel(a[:, 0], b[:, 0]) < TOL and _rel(a[:, 1:], b[:, 1:]) < TOL


@pytest.mark.parametrize("P,prec,images", [(10, 0, 0), (8, 1, 0), (6, 0, 2)])
def test_graph_replays_equal_eager(P, prec, images):
    import exafmm_amd as F
    print("graph P=%d precision=%d images=%d" % (P, prec, images), flush=True)
    L = 2 * np.pi
    x, q = F.make_bodies(20000, "cube1", 3)
    if images:
        x = x * (L / 2 * 0.999)
    kw = dict(images=images, cycle=L) if images else {}
    with F.FMM(P=P, precision=prec, **kw) as f:
        f.set_bodies(x, q)
        f.evaluate()
        eager = f.targets()
        if images == 0 and prec == 0:
            t = O.Tree(x, 64)
            ref = t.fmm(O.Kernel(P), x[t.perm], q[t.perm], 0.4, 5000)[np.argsort(t.perm)]
            assert _rel(eager[:, 0], ref[:, 0]) < 1e-10 and _rel(eager[:, 1:], ref[:, 1:]) < 1e-10
        for k in range(3):
            f.evaluate(graph=True)
            got = f.targets()
            assert _close(got, eager), k
            f.evaluate()  # eager work between replays
            f.direct(x[:10])
            f.synchronize()
        # new charges in place: the same graph solves them
        q2 = np.random.default_rng(7).uniform(-1, 1, len(q))
        f.set_charges(q2)
        f.evaluate()
        eager2 = f.targets()
        f.evaluate(graph=True)
        assert _close(f.targets(), eager2)
        # a new system: captured again
        x3, q3 = F.make_bodies(12000, "plummer1", 4)
        if images:
            x3 = np.clip(x3 / 10, -1, 1) * (L / 2 * 0.999)
        f.set_bodies(x3, q3)
        f.evaluate(graph=True)
        g3 = f.targets()
        f.evaluate()
        assert _close(g3, f.targets())
        f.evaluate(graph=True)
        assert _close(f.targets(), g3)
