"""On-disk formats: the replay record of state_io.py and MEDEA floe records
(synthetic data; no pickle is read or written)."""
import json

import numpy as np
import pytest

from icedem import scenario, state_io

# the keys of the reference's record, state_io.py:15-32
REFERENCE_KEYS = ['polygons', 'centroids', 'bonds', 'thicknesses', 'positions_history', 'rotations_history',
                  'bonds_history', 'ship_history', 'submergence_history', 'ship_length', 'ship_beam', 'ship_draft',
                  'n_frames', 'dt', 'output_2d', 'output_3d', 'simple_mode']


def _record(rng, n=7, nb=9, frames=5):
    polys = [rng.normal(0, 3, (int(rng.integers(3, 9)), 2)) for _ in range(n)]
    return dict(
        polygons=polys, centroids=np.array([p.mean(axis=0) for p in polys]),
        bonds=[(int(i), int(j)) for i, j in rng.integers(0, n, (nb, 2))], thicknesses=rng.uniform(0.5, 2, n),
        positions_history=[rng.normal(0, 5, (n, 2)).astype(np.float32) for _ in range(frames + 1)],
        rotations_history=[rng.normal(0, 1, n).astype(np.float32) for _ in range(frames + 1)],
        bonds_history=[rng.random(nb) < 0.8 for _ in range(frames + 1)],
        ship_history=[(rng.normal(0, 9, 2), float(rng.normal())) for _ in range(frames)],
        submergence_history=[rng.uniform(0, 1, n).astype(np.float32) for _ in range(frames + 1)],
        ship_length=60.0, ship_beam=12.0, ship_draft=5.0, n_frames=frames, dt=0.005,
        output_2d="out/anim.mp4", output_3d=None, simple_mode=True)


def test_state_round_trip_keeps_the_reference_keys_and_shapes(tmp_path):
    rec = _record(np.random.default_rng(0))
    path = tmp_path / "state.npz"
    state_io.save_simulation_state(filepath=path, **rec)
    z = np.load(path, allow_pickle=False)          # plain arrays only
    assert z["positions_history"].shape == (6, 7, 2)
    got = state_io.load_simulation_state(path)
    assert sorted(got) == sorted(REFERENCE_KEYS) == sorted(state_io.STATE_KEYS)
    for a, b in zip(got["polygons"], rec["polygons"]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(got["centroids"], rec["centroids"])
    assert got["bonds"] == rec["bonds"]
    for key in ("positions_history", "rotations_history", "bonds_history", "submergence_history"):
        assert len(got[key]) == len(rec[key])
        for a, b in zip(got[key], rec[key]):
            np.testing.assert_array_equal(a, b)
    for (pa, ha), (pb, hb) in zip(got["ship_history"], rec["ship_history"]):
        np.testing.assert_array_equal(pa, pb)
        assert ha == hb
    for key in ("ship_length", "ship_beam", "ship_draft", "n_frames", "dt", "output_2d", "output_3d", "simple_mode"):
        assert got[key] == rec[key]


def _square(cx, cy, s, cw=False):
    v = np.array([[cx - s, cy - s], [cx + s, cy - s], [cx + s, cy + s], [cx - s, cy + s]], float)
    return v[::-1].copy() if cw else v


def _records():
    sq = [_square(100, 200, 20), _square(170, 230, 15, cw=True), _square(260, 180, 25)]
    return [{"vertices": p.tolist(), "centre": p.mean(axis=0).tolist(), "area": scenario.polygon_area(p)} for p in sq]


def test_medea_records_and_files(tmp_path):
    recs = _records()
    polys, cents, meta = state_io.medea_floes_from_records(recs)
    assert meta["n_floes"] == 3 and meta["source"] == "MEDEA"
    np.testing.assert_allclose(meta["areas"], [1600.0, 900.0, 2500.0])
    np.testing.assert_array_equal(cents[0], [100.0, 200.0])
    for suffix in (".json", ".npz"):
        p = tmp_path / f"floes{suffix}"
        state_io.save_medea_floes(recs, p)
        p2, c2, m2 = state_io.load_medea_floes(p)
        for a, b in zip(p2, polys):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(c2, cents)
        np.testing.assert_array_equal(m2["areas"], meta["areas"])
    json.loads((tmp_path / "floes.json").read_text())
    with pytest.raises(ValueError):
        state_io.load_medea_floes(tmp_path / "floes.pkl")


def test_scale_floes_to_world_data_py_semantics():
    polys, cents, _ = state_io.medea_floes_from_records(_records())
    sp, sc = state_io.scale_floes_to_world(polys, cents, 0.5, center_domain=False)
    np.testing.assert_array_equal(sp[0], polys[0] * 0.5)
    np.testing.assert_array_equal(sc, cents * 0.5)
    sp, sc = state_io.scale_floes_to_world(polys, cents, 0.5)
    allv = np.vstack(sp)
    np.testing.assert_allclose(allv.min(axis=0) + allv.max(axis=0), 0.0, atol=1e-12)
    # bounding box of the vertices [80, 285] x [155, 245] -> centre (182.5, 200) px
    np.testing.assert_allclose(sc[0], (np.array([100.0, 200.0]) - [182.5, 200.0]) * 0.5)
    sp2, sc2 = state_io.scale_floes_to_world(polys, cents, 0.5, subsample_region=(-60.0, -20.0, -50.0, 50.0))
    assert len(sp2) == 1 and np.allclose(sc2[0], sc[0])
    sp3, _ = state_io.scale_floes_to_world(polys, cents, 0.5, subsample_region=(1e3, 2e3, 1e3, 2e3))
    assert len(sp3) == 3                            # nothing inside: every floe kept


def test_scene_from_medea_tessellates_and_bonds():
    sc = state_io.scene_from_medea(_records(), meters_per_pixel=1.0, thickness=[1.0, 2.0, 0.5], cells_per_floe=12)
    assert sc.n_floes == 3
    areas = np.array([scenario.polygon_area(p) for p in sc.polygons])
    for f, want in enumerate((1600.0, 900.0, 2500.0)):
        sel = np.asarray(sc.parent_ids) == f
        assert sel.sum() >= 4
        assert abs(areas[sel].sum() - want) < 1e-6 * want
        np.testing.assert_array_equal(sc.thicknesses[sel], [1.0, 2.0, 0.5][f])
    # every cell is counter-clockwise and its centroid is the vertex mean
    for p, c in zip(sc.polygons, sc.centroids):
        a2 = np.sum(p[:, 0] * np.roll(p[:, 1], -1) - np.roll(p[:, 0], -1) * p[:, 1])
        assert a2 > 0
        np.testing.assert_allclose(c, p.mean(axis=0))
    # bonds: exactly the pairs with |c_i - c_j| < 1.5 (r_i + r_j), r = sqrt(A / pi) (bonds.py:6-16)
    r = np.sqrt(areas / np.pi)
    c = np.asarray(sc.centroids)
    d = np.linalg.norm(c[:, None] - c[None], axis=2)
    want = [(int(i), int(j)) for i, j in zip(*np.nonzero(np.triu(d < 1.5 * (r[:, None] + r[None]), 1)))]
    assert sorted(sc.bonds) == sorted(want)
