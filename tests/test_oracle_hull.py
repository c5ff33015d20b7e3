"""Hull-ice narrow phase of the oracle (a4-a5): special cases, brute force and
the reference's quirks (jax_geometry.py): bow at -x with a +0.5 offset, the
closing edge not sampled when a polygon is padded, first-index tie-breaks."""
import math

import numpy as np

from oracle import model, step


def _oracle(polys, ft=np.float64, L=40.0, B=10.0, K=3, hull=None):
    n = len(polys)
    cen = np.array([np.mean(p, axis=0) for p in polys])
    cfg = model.OracleConfig(n_edge_samples=K)
    st = model.build_static_data(polys, cen, [], np.ones(n), np.arange(n), cfg)
    sd = model.build_ship_dynamics_data(1e6, L, B, hull)
    f = model.Fossen(1e6, L)
    return step.Oracle(st, cfg, sd, L, B, 1, 0.05, 0.05, False, f.B, f.T_n, dtype=ft), cen


def test_point_at_hull_centre_and_outside():
    o, _ = _oracle([np.array([[0, 0], [1, 0], [0, 1.0]])])
    pen, nx, ny = o.point_to_ship(np.array([0.0, 100.0]), np.array([0.0, 0.0]), 0.0, 0.0, 0.0)
    assert pen[0] == min(20.0, 5.0) and pen[1] == 0.0 and nx[1] == 0 and ny[1] == 0
    # side contact at rel_y = 3: penetration hb - 3, outward normal +y
    pen, nx, ny = o.point_to_ship(np.array([5.0]), np.array([3.0]), 0.0, 0.0, 0.0)
    assert pen[0] == 2.0 and nx[0] == 0.0 and ny[0] == 1.0
    # rotated ship: world normal rotated by the heading
    pen, nx, ny = o.point_to_ship(np.array([-3.0]), np.array([5.0]), 0.0, 0.0, math.pi / 2)
    assert abs(pen[0] - 2.0) < 1e-12 and abs(nx[0] - (-1.0)) < 1e-12 and abs(ny[0]) < 1e-12


def test_bow_is_at_negative_x_with_offset():
    o, _ = _oracle([np.array([[0, 0], [1, 0], [0, 1.0]])])
    hl, hb = 20.0, 5.0
    x = -0.9 * hl
    pen, nx, ny = o.point_to_ship(np.array([x]), np.array([0.5]), 0.0, 0.0, 0.0)
    t = ((0.9) - 0.6) / 0.4
    lhb = max(hb * (1 - 0.9 * t ** 1.5), hb * 0.05)
    expected = min(x + hl, lhb - 0.5) + 0.5
    assert abs(pen[0] - expected) < 1e-12
    ang = 0.3 + 0.5 * t
    assert abs(nx[0] + math.cos(ang)) < 1e-12 and abs(ny[0] - math.sin(ang)) < 1e-12
    # +x end of the ship is a flat, untapered face
    pen2, _, _ = o.point_to_ship(np.array([0.9 * hl]), np.array([4.9]), 0.0, 0.0, 0.0)
    assert abs(pen2[0] - 0.1) < 1e-9


def test_hull_point_inside_square_pass2():
    sq = np.array([[-1.0, -1.0], [1.0, -1.0], [1.0, 1.0], [-1.0, 1.0]]) + [30.0, 0.3]
    hull = np.array([[10.0, 0.0], [-10.0, 2.0], [-10.0, -2.0]])   # only (10,0) ends inside the square
    o, cen = _oracle([sq], hull=hull)
    wx, wy = o.world_polys(cen, np.zeros(1))
    pen, nx, ny, cx, cy = o.pass2(wx, wy, 20.0, 0.0, 0.0)
    assert abs(pen[0] - 0.7) < 1e-9           # nearest edge y = -0.7 (relative to the hull point)
    assert nx[0] == 0.0 and ny[0] == -1.0 and cx[0] == 30.0 and cy[0] == 0.0


def test_closing_edge_not_sampled_when_padded():
    """jax_geometry.py:194,209 + padding repeats the last vertex: the edge
    v_{nv-1} -> v_0 is sampled only when nv == max_V."""
    tri = np.array([[-30.0, 20.0], [30.0, 20.0], [0.0, 25.0]])
    tri = np.array([[-30.0, -2.0], [-25.0, 20.0], [30.0, 2.0]])
    quad = np.array([[100.0, 100.0], [101.0, 100.0], [101.0, 101.0], [100.0, 101.0]])
    o_pad, cen_pad = _oracle([tri, quad], L=20.0, B=4.0, K=4)
    o_full, cen_full = _oracle([tri], L=20.0, B=4.0, K=4)
    # ship sits on the closing edge (30,2) -> (-30,-2) and away from the vertices
    for o, cen in ((o_pad, cen_pad), (o_full, cen_full)):
        wx, wy = o.world_polys(cen, np.zeros(len(cen)))
        p1 = o.pass1(wx, wy, 0.0, 0.0, 0.0)
        if o is o_pad:
            pen_pad = p1[0][0]
        else:
            pen_full = p1[0][0]
    assert pen_pad == 0.0 and pen_full > 0.0


def test_first_index_tie_break():
    o, _ = _oracle([np.array([[0, 0], [1, 0], [0, 1.0]])])
    px = np.array([[5.0, 5.0, 7.0]]); py = np.array([[3.0, -3.0, 3.0]])
    pen, nx, ny = o.point_to_ship(px, py, 0.0, 0.0, 0.0)
    assert pen[0, 0] == pen[0, 1] == pen[0, 2]
    best = np.argmax(pen, axis=1)
    assert best[0] == 0 and ny[0, best[0]] == 1.0


def test_pass1_brute_force():
    rng = np.random.default_rng(0)
    polys = [np.array([[math.cos(a), math.sin(a)] for a in np.sort(rng.uniform(0, 2 * np.pi, 6))]) * 3
             + rng.uniform(-25, 25, 2) for _ in range(30)]
    o, cen = _oracle(polys, K=3)
    rot = rng.uniform(-3, 3, len(polys))
    wx, wy = o.world_polys(cen, rot)
    pen, nx, ny, cx, cy = o.pass1(wx, wy, 1.0, -2.0, 0.3)
    V = wx.shape[1]
    t = step.edge_sample_t(3, np.float64)
    for e in range(len(polys)):
        best = (-1.0, None)
        for k in range(V):
            pts = [(wx[e, k], wy[e, k])]
            kn = (k + 1) % V
            pts += [(wx[e, k] * (1 - tt) + wx[e, kn] * tt, wy[e, k] * (1 - tt) + wy[e, kn] * tt) for tt in t]
            for (x, y) in pts:
                p, a, b = o.point_to_ship(np.array([x]), np.array([y]), 1.0, -2.0, 0.3)
                pv = p[0] if o.poly_mask[e, k] else 0.0
                if pv > best[0]:
                    best = (pv, (a[0], b[0]))
        assert abs(pen[e] - best[0]) < 1e-12 and (best[0] == 0 or (nx[e], ny[e]) == best[1])
