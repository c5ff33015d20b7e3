"""Connected components and COM velocities (a16), the reference's own cases
(test_jax_full_scan.py:65-182) plus BFS on random graphs."""
import numpy as np

from oracle import bonds


def _labels(n, edges, intact=None):
    bi = np.array([a for a, b in edges], np.int32); bj = np.array([b for a, b in edges], np.int32)
    it = np.ones(len(edges), bool) if intact is None else np.asarray(intact, bool)
    return bonds.find_connected_components(n, bi, bj, it)


def _partition(labels):
    g = {}
    for i, l in enumerate(labels):
        g.setdefault(int(l), []).append(i)
    return {min(v): sorted(v) for v in g.values()}


def test_reference_cases():
    assert np.all(_labels(4, [(0, 1), (1, 2), (2, 3)]) == 0)
    assert np.all(_labels(4, [(0, 1), (0, 2), (0, 3)]) == 0)
    l = _labels(4, [(0, 1), (2, 3)])
    assert l[0] == l[1] == 0 and l[2] == l[3] == 2
    assert list(_labels(1, [])) == [0]
    assert list(_labels(3, [(0, 1), (1, 2)], [False, False])) == [0, 1, 2]
    assert np.all(_labels(20, [(i, i + 1) for i in range(19)]) == 0)


def test_random_graphs_match_bfs():
    rng = np.random.default_rng(0)
    for trial in range(20):
        n = int(rng.integers(2, 200))
        m = int(rng.integers(0, 2 * n))
        e = rng.integers(0, n, (m, 2))
        e = [tuple(map(int, x)) for x in e if x[0] != x[1]]
        it = rng.random(len(e)) < 0.7
        bi = np.array([a for a, b in e], np.int32); bj = np.array([b for a, b in e], np.int32)
        lab = bonds.find_connected_components(n, bi, bj, it, n_iter=n)
        comps = bonds.bfs_components(n, e, it)
        assert _partition(lab) == {min(c): sorted(c) for c in comps}


def test_label_propagation_needs_diameter_iterations():
    e = [(i, i + 1) for i in range(99)]
    bi = np.array([a for a, b in e], np.int32); bj = np.array([b for a, b in e], np.int32)
    lab = bonds.find_connected_components(100, bi, bj, np.ones(99, bool), n_iter=64)
    assert lab[0] == 0 and lab[99] == 99 - 64   # one hop per sweep


def test_com_velocities():
    v = bonds.component_com_velocities(np.array([0, 0, 2, 2]), np.array([1.0, 1.0, 2.0, 2.0]),
                                       np.array([[1.0, 0.0], [3.0, 0.0], [0.0, 1.0], [0.0, 3.0]]))
    np.testing.assert_allclose(v, [[2, 0], [2, 0], [0, 2], [0, 2]], atol=1e-10)
