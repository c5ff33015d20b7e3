"""The planner's two scorers besides the DEM, and what the network gains.

* The reference planners' KE costmap (build_costmap_jax, jax_costmap.py:
  98-330): closed forms of one element (peak = the KE constant, zero beyond
  the effective radius), maximum (not sum) where elements overlap, the
  open-water floor, the square max-window dilation with edge padding.
* Ship-only rollouts (the costmap planners' rollout): the same ship model as
  the DEM's ship_mode 1 -- with the ice far away, the DEM rollout's KE loss;
  on a constant costmap the integral is that constant times the path length.
* The gain test: on held-out DEM rollouts (random starts and controls over
  the rollout patches), the trained network's cost error against the
  costmap cost calibrated on the same training rollouts.
"""
import math
import types

import numpy as np
import pytest
import torch

from icedem import planner as pl, rollouts as ro, scenario
from icedem.properties import ShipProperties
from icedem.surrogate import Arch, Surrogate

pytestmark = pytest.mark.gpu


def _ke(m, M, v, w=5.0, s=1e-6):
    return min(max(w * v * v * m * M * (m + 2 * M) / (2 * (M + m) ** 2) * s, 0.0), 1e8)


def test_ke_costmap_closed_forms():
    from icedem import _lib
    lib = _lib.load()
    dev = "cuda"
    # elements at (5, 5) r 1 m 100 t, (5.4, 5) r 1 m 400 t, (20, 20) r 2 m 800 t on a 1 m grid from (0, 0)
    pos = torch.tensor([5.0, 5.4, 20.0, 5.0, 5.0, 20.0], device=dev)
    rad = torch.tensor([1.0, 1.0, 2.0], device=dev)
    mass = torch.tensor([1e5, 4e5, 8e5], device=dev)
    rows, cols, res, M, v = 30, 32, 1.0, 2.2e6, 2.0
    work = torch.empty(rows * cols, dtype=torch.int32, device=dev)

    def build(half, base=0.1, n=3):
        out = torch.empty(rows * cols, dtype=torch.float32, device=dev)
        assert lib.icesur_ke_costmap(pl._ptr(pos), pl._ptr(rad), pl._ptr(mass), n, rows, cols, 0.0, 0.0, res, 5.0, v,
                                     M, 1e-6, 0.1, base, half, pl._ptr(work), pl._ptr(out), None) == 0
        torch.cuda.synchronize()
        return out.view(rows, cols).cpu().numpy().astype(np.float64)

    g = build(0)
    cy, cx = np.meshgrid(np.arange(rows) + 0.5, np.arange(cols) + 0.5, indexing="ij")
    ref = np.full((rows, cols), 0.1)
    for (x, y, r, m) in [(5.0, 5.0, 1.0, 1e5), (5.4, 5.0, 1.0, 4e5), (20.0, 20.0, 2.0, 8e5)]:
        er = r * 1.1 + 0.5
        f = np.maximum(0.0, (er * er - ((cx - x) ** 2 + (cy - y) ** 2)) / (er * er))
        ref = np.maximum(ref, _ke(m, M, v) * f)
    np.testing.assert_allclose(g, ref, rtol=2e-6, atol=1e-5)
    # the heavier of two overlapping elements wins a shared cell (maximum, not sum)
    k1, k2 = _ke(1e5, M, v), _ke(4e5, M, v)
    assert k2 > k1 > 0.5
    f1 = 1.0 - ((5.5 - 5.0) ** 2 + 0.25) / 1.6 ** 2
    f2 = 1.0 - ((5.5 - 5.4) ** 2 + 0.25) / 1.6 ** 2
    assert g[5, 5] == pytest.approx(max(k1 * f1, k2 * f2), rel=1e-5) and g[5, 5] < k1 * f1 + k2 * f2
    # far from every element: the open-water floor
    assert g[29, 0] == pytest.approx(0.1)
    # dilation: each cell is the maximum of its (2h+1)^2 window (edge padding)
    h = 2
    gd = build(h)
    pad = np.pad(g, h, mode="edge")
    win = np.max(np.stack([pad[i:i + rows, j:j + cols] for i in range(2 * h + 1) for j in range(2 * h + 1)]), 0)
    np.testing.assert_array_equal(gd, win)


def _far_batch(n_inst=8):
    sc = scenario.scene_patch(1234, 120, 0.3)
    rb = ro.RolloutBatch([sc], [0] * n_inst)
    rb.set_mode("frames")
    # start 400 m from the ice: no contact within 400 frames
    rb.ship_init[:, 1] -= 400.0
    return rb


def test_ship_rollouts_match_dem_without_ice():
    rb = _far_batch()
    ctl = torch.as_tensor(ro.nominal_controls(rb.R, 40, seed=3))
    rb.reset()
    out = rb.run(ctl, 10, 400).cpu().numpy()
    assert np.all(out[:, 0] == 0.0)
    so = pl.ship_rollouts(rb.engines[0], rb.ship_init, ctl, 10, 400).cpu().numpy()
    np.testing.assert_allclose(so[:, 1], out[:, 1], rtol=1e-6, atol=1e-3)
    # a constant costmap integrates to the constant times the path length
    c = 0.75
    cm = types.SimpleNamespace(grid=torch.full((50 * 60,), c, device="cuda"), rows=50, cols=60, x0=-1e4, y0=-1e4,
                               res=1000.0, spec=types.SimpleNamespace(oob_penalty=0.0))
    so2 = pl.ship_rollouts(rb.engines[0], rb.ship_init, ctl, 10, 400, cm).cpu().numpy()
    np.testing.assert_allclose(so2[:, 0], c * so2[:, 2], rtol=1e-5)
    np.testing.assert_array_equal(so2[:, 1:], so[:, 1:])
    # straight ahead (no yaw command, no initial yaw rate): path length = distance travelled
    straight = torch.zeros((rb.R, 40, 2))
    straight[:, :, 0] = 0.5
    rb.reset()
    rb.run(straight, 10, 400)
    s3, snap = pl.ship_rollouts(rb.engines[0], rb.ship_init, straight, 10, 400, None, 4)
    s3, snap = s3.cpu().numpy(), snap.cpu().numpy()
    d = (rb.ships[:, :2] - rb.ship_init[:, :2]).norm(dim=1).cpu().numpy()
    np.testing.assert_allclose(s3[:, 2], d, rtol=1e-3)
    # the last snapshot is the DEM rollout's end pose seen from the start pose
    psi0 = rb.ship_init[:, 2].cpu().numpy()
    e = (rb.ships[:, :2] - rb.ship_init[:, :2]).cpu().numpy()
    along = np.cos(psi0) * e[:, 0] + np.sin(psi0) * e[:, 1]
    np.testing.assert_allclose(snap[:, 3, 0], along, rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(snap[:, 3, 1], 0.0, atol=1e-3)
    np.testing.assert_allclose(snap[:, :, 2], 0.0, atol=1e-6)
    assert np.all(np.diff(snap[:, :, 0], axis=1) > 0)
    rb.close()


def test_network_gains_over_costmap_cost():
    """Same rollouts, same split: the trained network's median relative cost
    error is below the calibrated costmap cost's, and it ranks the held-out
    rollouts better."""
    scenes = ro.rollout_patches(128, workers=8)
    rb = ro.RolloutBatch(scenes, list(range(128)))
    rb.set_mode("frames")
    data = pl.make_dataset(rb, 80, seed=7)
    rb.close()
    N = data["raster"].shape[0]
    w, cm_err = pl.costmap_baseline(data)
    net = Surrogate(Arch(), seed=0)
    norm, hist = pl.train(net, data, epochs=20, batch=512, lr=1e-3)
    sur_err = hist[-1]["val_rel_cost_err"]
    val, _ = pl.split_indices(N)
    cost = pl.cost_of_targets(data["target"].cpu().numpy())[val]
    ii = torch.as_tensor(val, device="cuda")
    ps = pl.cost_of_targets(norm.inv(net.predict(data["raster"][ii], data["extra"][ii])).cpu().numpy())
    pc = data["costmap_feat"].cpu().numpy()[val] @ w

    def sp(a, b):
        ra, rb_ = np.argsort(np.argsort(a)), np.argsort(np.argsort(b))
        return float(np.corrcoef(ra, rb_)[0, 1])
    print(f"{N} rollouts: surrogate val rel cost err {sur_err}, spearman {sp(ps, cost):.3f}; "
          f"costmap cost {cm_err}, spearman {sp(pc, cost):.3f}; weights {w}", flush=True)
    assert sur_err["median"] < cm_err["median"]
    assert sp(ps, cost) > sp(pc, cost)


def test_planner_reports_dem_costs_and_widens_its_judge():
    """Every reported cost is a DEM cost of the reported controls; the
    fallback judges more candidates, and the next iteration judges its whole
    budget in one launch."""
    sc = scenario.scene_patch(5000, 200, 0.4)
    tmp = ro.RolloutBatch([sc], [0])
    start = tmp.ship_init[0].clone()
    tmp.close()
    net = Surrogate(Arch(), seed=0)              # untrained: its error forces the fallback
    norm = pl.Norm(np.zeros(3, np.float32), np.ones(3, np.float32))
    P = pl.Planner(net, norm, sc, start, k=64, max_judged=256, err_threshold=0.05)
    nominal = np.tile(np.array([0.7, 0.0], np.float32), (40, 1))
    r1 = P.iterate(nominal, S=2048, seed=1)
    assert r1["judged"] == 256 and P.width == 256
    again = P.dem_costs(torch.as_tensor(r1["best_controls"][None], device="cuda"))
    assert again[0] == pytest.approx(r1["best_cost"], rel=0, abs=0)
    r2 = P.iterate(nominal, S=2048, seed=2)
    assert r2["judged"] == 256
    again = P.dem_costs(torch.as_tensor(r2["best_controls"][None], device="cuda"), 256)
    assert again[0] == r2["best_cost"]
    P.close()


def test_jitter_mean_is_the_mean_of_shifted_runs():
    """run_jittered: zero offsets give the plain batched run, a ring gives the
    float64 mean of the runs from the shifted starts, bit for bit."""
    scenes = ro.rollout_patches(4, workers=1)
    rb = ro.RolloutBatch(scenes, list(range(4)))
    rb.set_mode("frames")
    ctl = torch.as_tensor(ro.nominal_controls(4, 40, seed=3))
    st = rb.ship_init.clone()
    rb.reset()
    plain = rb.run(ctl, 10, 400).cpu().numpy().astype(np.float64)
    assert np.any(plain[:, 0] > 0), "the ship must reach the ice"
    np.testing.assert_array_equal(pl.run_jittered(rb, st, ctl, 10, 400, np.zeros((1, 2))), plain)
    offs = pl.jitter_offsets(3, 0.05)
    acc = None
    for o in offs:
        s2 = st.clone()
        s2[:, 0] += float(o[0])
        s2[:, 1] += float(o[1])
        rb.ship_init.copy_(s2)
        rb.reset()
        r = rb.run(ctl, 10, 400).cpu().numpy().astype(np.float64)
        acc = r if acc is None else acc + r
    np.testing.assert_array_equal(pl.run_jittered(rb, st, ctl, 10, 400, offs), acc / 3)
    assert not np.array_equal(acc / 3, plain)
    rb.close()


def test_costmap_ranked_iteration_reports_dem_costs():
    """The planner ranked by the calibrated costmap instead of the network:
    the same DEM judge, so the reported best cost is the DEM cost of the
    reported controls; an unknown ranker is refused."""
    sc = scenario.scene_patch(5001, 200, 0.4)
    tmp = ro.RolloutBatch([sc], [0])
    start = tmp.ship_init[0].clone()
    tmp.close()
    net = Surrogate(Arch(), seed=0)
    norm = pl.Norm(np.zeros(3, np.float32), np.ones(3, np.float32))
    P = pl.Planner(net, norm, sc, start, k=64, costmap_weights=[1.0, 0.5, 2.0])
    nominal = np.tile(np.array([0.7, 0.0], np.float32), (40, 1))
    r = P.iterate(nominal, S=2048, seed=4, ranker="costmap")
    again = P.dem_costs(torch.as_tensor(r["best_controls"][None], device="cuda"), P.width)
    assert again[0] == r["best_cost"]
    with pytest.raises(ValueError):
        P.iterate(nominal, S=256, ranker="random")
    P.close()


def test_mppi_update_matches_the_weighted_mean():
    """run_mppi's update (jax_mppi_planner.py:95-101) against float64 NumPy."""
    rng = np.random.default_rng(4)
    for S, D, T in [(1000, 80, 0.1), (16384, 80, 0.5), (3, 7, 2.0)]:
        c = rng.normal(10.0, 1.0, S).astype(np.float32)
        u = np.clip(rng.normal(0.5, 0.3, (S, D)), -1, 1).astype(np.float32)
        got = pl.mppi_update(torch.as_tensor(c, device="cuda"), torch.as_tensor(u, device="cuda"), T).cpu().numpy()
        w = np.exp(-(c.astype(np.float64) - c.min()) / T)
        ref = np.clip((w[:, None] * u).sum(0) / w.sum(), -1, 1)
        np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-6)
    # a single dominant sample wins outright; clipping to [-1, 1]
    c = np.array([5.0, 0.0, 5.0], np.float32)
    u = np.array([[0.0, 0.0], [3.0, -0.25], [0.0, 0.0]], np.float32)
    got = pl.mppi_update(torch.as_tensor(c, device="cuda"), torch.as_tensor(u, device="cuda"), 0.01).cpu().numpy()
    np.testing.assert_allclose(got, [1.0, -0.25], atol=1e-6)


def test_mppi_loop_runs_on_planner_scores():
    sc = scenario.scene_patch(5001, 200, 0.4)
    tmp = ro.RolloutBatch([sc], [0])
    start = tmp.ship_init[0].clone()
    tmp.close()
    net = Surrogate(Arch(), seed=0)
    norm = pl.Norm(np.zeros(3, np.float32), np.ones(3, np.float32))
    P = pl.Planner(net, norm, sc, start, k=64, max_judged=64)
    nominal = np.tile(np.array([0.7, 0.0], np.float32), (40, 1))
    r = P.mppi(nominal, iterations=2, temperature=0.1, S=1024, seed=3)
    assert r["nominal"].shape == (40, 2) and np.all(np.abs(r["nominal"]) <= 1.0)
    assert len(r["history"]) == 2 and np.isfinite(r["dem_cost"])
    again = P.dem_costs(torch.as_tensor(r["nominal"][None], device="cuda"))[0]
    assert again == r["dem_cost"]
    P.close()


def _mbd_ref(c, u, T, ybar, ab, a, abp):
    """run_mbd's reverse step (jax_mbd_planner.py:1363-1396) in float64 NumPy."""
    c = c.astype(np.float64); u = u.astype(np.float64); yb = ybar.astype(np.float64)
    logp = (-c - (-c).max()) / T
    w = np.exp(logp) / np.exp(logp).sum()
    ynew = (w[:, None] * u).sum(0)
    yi = yb * np.sqrt(ab)
    score = (-yi + np.sqrt(ab) * ynew) / (1.0 - ab)
    yim1 = (yi + (1.0 - ab) * score) / np.sqrt(a)
    return np.clip(yim1 / np.sqrt(abp), -1.0, 1.0)


def test_mbd_update_matches_float64():
    """icesur_mbd_update against float64 NumPy of the same step, on the
    reference's schedule (beta 1e-5 .. 0.05, 50 steps) and on arbitrary
    schedule values (where the mean does not cancel)."""
    rng = np.random.default_rng(6)
    _, al, ab, _ = pl.mbd_schedule(50, 1e-5, 0.05)
    cases = [(2048, 80, 0.01, ab[i], al[i], ab[i - 1]) for i in (1, 10, 49)]
    cases += [(1000, 80, 0.1, 0.9, 0.95, 0.7), (5, 3, 1.0, 0.3, 0.6, 0.5)]
    for S, D, T, ab_i, a_i, abp in cases:
        c = rng.normal(10.0, 1.0, S).astype(np.float32)
        u = np.clip(rng.normal(0.3, 0.4, (S, D)), -1, 1).astype(np.float32)
        yb = np.clip(rng.normal(0.2, 0.5, D), -1, 1).astype(np.float32)
        got = pl.mbd_update(torch.as_tensor(c, device="cuda"), torch.as_tensor(u, device="cuda"), T,
                            torch.as_tensor(yb, device="cuda"), ab_i, a_i, abp).cpu().numpy()
        ref = _mbd_ref(c, u, T, yb, float(ab_i), float(a_i), float(abp))
        np.testing.assert_allclose(got, ref, rtol=2e-5, atol=5e-6)


def test_mbd_step_on_its_schedule_is_the_weighted_mean():
    """With alphas_bar[i] = alphas[i] alphas_bar[i-1] the score step cancels:
    Ybar_{i-1} = clip(sum w Y0s / sum w), whatever Ybar_i was -- the MPPI
    update of the same costs and samples."""
    rng = np.random.default_rng(8)
    _, al, ab, _ = pl.mbd_schedule(50, 1e-5, 0.05)
    for i in (1, 25, 49):
        c = torch.as_tensor(rng.normal(5.0, 2.0, 4096).astype(np.float32), device="cuda")
        u = torch.as_tensor(np.clip(rng.normal(0.0, 0.6, (4096, 40, 2)), -1, 1).astype(np.float32), device="cuda")
        yb = torch.as_tensor(rng.uniform(-1, 1, (40, 2)).astype(np.float32), device="cuda")
        got = pl.mbd_update(c, u, 0.05, yb, ab[i], al[i], ab[i - 1]).cpu().numpy()
        ref = pl.mppi_update(c, u, 0.05).cpu().numpy()
        np.testing.assert_allclose(got, ref, atol=2e-5)


def test_mbd_loop_runs_on_planner_scores():
    sc = scenario.scene_patch(5001, 200, 0.4)
    tmp = ro.RolloutBatch([sc], [0])
    start = tmp.ship_init[0].clone()
    tmp.close()
    net = Surrogate(Arch(), seed=0)
    norm = pl.Norm(np.zeros(3, np.float32), np.ones(3, np.float32))
    P = pl.Planner(net, norm, sc, start, k=64, max_judged=64)
    r = P.mbd(None, n_diffuse=3, temperature=0.1, S=1024, seed=3)
    assert r["ybar"].shape == (40, 2) and np.all(np.abs(r["ybar"]) <= 1.0)
    assert [h["i"] for h in r["history"]] == [2, 1] and np.isfinite(r["dem_cost"])
    assert r["history"][0]["sigma"] > r["history"][1]["sigma"]
    again = P.dem_costs(torch.as_tensor(r["ybar"][None], device="cuda"))[0]
    assert again == r["dem_cost"]
    P.close()
