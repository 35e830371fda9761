"""The request loop (fmm_server_*): every result is the same alone and in a
batch, batching and slot overlap actually happen, and memory admission holds.

Alone vs in the mix: 1e-12 relative (L2 per request).  A system's cells,
lists, M2L tiles and chunks are identical in both cases; only the order of
FP64 atomic additions into its local expansions may differ.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX_TOL = 1e-12


def _F():
    import exafmm_amd
    return exafmm_amd


def _rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def _systems(F, k, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(2 ** rng.uniform(9, 14))
        x, q = F.make_bodies(n, "cube1" if i % 2 == 0 else "plummer1", 100 + i)
        out.append((x, q))
    return out


def _serve(F, systems, precision, m2l_mode=0, **cfg):
    s = F.Server(P=8, precision=precision, m2l_mode=m2l_mode, **cfg)
    ids = [s.submit(x, q) for x, q in systems]
    res = [s.wait(i) for i in ids]
    st = s.stats()
    s.close()
    return res, st


# mode 0 (AUTO) runs tiles alone (one request per batch) and grouped chunks
# in the mix: FP64 only, where the two agree to the order of the sums.  Mixed
# precision on explicit tiles is not batch-invariant (a merged forest's
# by-level groups leave pairs of other radii to the per-target kernel), which
# is why AUTO takes chunks for batches of several requests.
@pytest.mark.parametrize("precision,mode", [("fp64", 0), ("fp64", 3), ("mixed", 2), ("fp64", 2)])
def test_results_equal_alone_and_in_mix(precision, mode):
    F = _F()
    prec = F.FMM_MIXED if precision == "mixed" else F.FMM_FP64
    systems = _systems(F, 12, 1)
    alone, st1 = _serve(F, systems, prec, slots=1, batch_requests=1, m2l_mode=mode)
    mix, st2 = _serve(F, systems, prec, slots=3, batch_requests=64, batch_delay_ms=500.0, m2l_mode=mode)
    assert st1["max_batch_requests"] == 1 and st1["batches"] == 12
    assert st2["max_batch_requests"] > 1 and st2["requests_done"] == 12
    for (ta, ia), (tm, im), (x, q) in zip(alone, mix, systems):
        assert ta.shape == (len(x), 4)
        assert _rel(tm[:, 0], ta[:, 0]) < MIX_TOL and _rel(tm[:, 1:], ta[:, 1:]) < MIX_TOL
    # and equal to a context solving the system by itself
    for (ta, _), (x, q) in zip(alone[:3], systems[:3]):
        f = F.FMM(P=8, precision=prec, m2l_mode=mode)
        f.set_bodies(x, q)
        f.evaluate()
        t = f.targets(input_order=True)
        f.close()
        assert _rel(ta[:, 0], t[:, 0]) < MIX_TOL and _rel(ta[:, 1:], t[:, 1:]) < MIX_TOL


def test_plans_built_on_the_device():
    """plan_on_device: the builders build trees and lists on the device (the
    same trees and pairs, lists in source order): results equal the
    host-planned loop's to the rounding of the summation order."""
    F = _F()
    systems = _systems(F, 12, 2)
    host, _ = _serve(F, systems, F.FMM_FP64, slots=3, batch_requests=64)
    dev, st = _serve(F, systems, F.FMM_FP64, slots=3, batch_requests=64, plan_on_device=1)
    assert st["requests_done"] == 12
    for (th, _), (td, _) in zip(host, dev):
        assert _rel(td[:, 0], th[:, 0]) < MIX_TOL and _rel(td[:, 1:], th[:, 1:]) < MIX_TOL


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_planned_tiles_one_request_per_batch(precision):
    """plan_on_device with M2L tiles and one request per batch: each request's
    tree, lists, M2L groups and tiles are built on the device and uploaded
    device to device.  Same results as the host-planned loop with tiles (FP64:
    1e-12; mixed: the FP32 P2P sums each target's sources in list order,
    which differs between device and host lists, so 1e-6 there) and equal
    to a context solving the request alone from its own device plan (1e-12)."""
    F = _F()
    prec = F.FMM_MIXED if precision == "mixed" else F.FMM_FP64
    systems = _systems(F, 10, 3)
    host, _ = _serve(F, systems, prec, m2l_mode=F.M2L_TILES, slots=2, batch_requests=1)
    dev, st = _serve(F, systems, prec, m2l_mode=F.M2L_TILES, slots=3, batch_requests=1, plan_on_device=1)
    assert st["requests_done"] == 10 and st["max_batch_requests"] == 1
    tol = MIX_TOL if precision == "fp64" else 1e-6
    for (th, _), (td, _) in zip(host, dev):
        assert _rel(td[:, 0], th[:, 0]) < tol and _rel(td[:, 1:], th[:, 1:]) < tol
    for (td, _), (x, q) in zip(dev[:3], systems[:3]):
        f = F.FMM(P=8, precision=prec, m2l_mode=F.M2L_TILES)
        f.set_bodies_plan(F.Plan(x, lists_on=f), q)
        f.evaluate()
        t = f.targets(input_order=True)
        f.close()
        assert _rel(td[:, 0], t[:, 0]) < MIX_TOL and _rel(td[:, 1:], t[:, 1:]) < MIX_TOL


def test_memory_admission_serialises_batches():
    """A budget below any footprint admits a batch only when none is in
    flight: batch intervals never overlap."""
    F = _F()
    systems = _systems(F, 8, 2)
    res, st = _serve(F, systems, F.FMM_FP64, slots=3, batch_requests=2, device_budget=1)
    spans = sorted({(i["t_started_ms"], i["t_done_ms"]) for _, i in res})
    for (a0, a1), (b0, b1) in zip(spans, spans[1:]):
        assert b0 >= a1
    assert st["requests_done"] == 8


def test_empty_and_tiny_requests():
    F = _F()
    s = F.Server(P=6, slots=2)
    ids = [s.submit(np.zeros((0, 3)), np.zeros(0)), s.submit(np.array([[0.1, 0.2, 0.3]]), np.array([1.0]))]
    t0, _ = s.wait(ids[0])
    t1, _ = s.wait(ids[1])
    s.close()
    assert t0.shape == (0, 4) and np.array_equal(t1, np.zeros((1, 4)))


def test_destroy_releases_blocked_waits():
    """Destroying the loop while requests are queued: a wait blocked on one
    of them returns (the result, or FMM_ERR_STATE if it never ran) before
    destroy returns, and destroy does not hang."""
    import threading
    import time
    F = _F()
    s = F.Server(P=8, slots=1, batch_requests=1)
    systems = _systems(F, 6, 3) + [F.make_bodies(1 << 17, "cube1", 9)]
    ids = [s.submit(x, q) for x, q in systems]
    out = {}

    def waiter():
        try:
            s.wait(ids[-1])
            out["rc"] = 0
        except F.FMMError as e:
            out["rc"] = str(e)

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.3)
    s.close()
    th.join(timeout=60)
    assert not th.is_alive() and "rc" in out
