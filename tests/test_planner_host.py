"""Host-side planner pieces (no GPU): run_mbd's noise schedule."""
import numpy as np

from icedem import planner as pl


def test_mbd_schedule_closed_forms():
    """jax_mbd_planner.py:1358-1361: betas evenly spaced from beta0 to betaT,
    alphas_bar the running product of 1 - beta, sigma = sqrt(1 - alphas_bar)."""
    betas, alphas, ab, sig = pl.mbd_schedule(50, 1e-5, 0.05)
    assert betas.dtype == alphas.dtype == ab.dtype == sig.dtype == np.float32
    np.testing.assert_allclose(betas[[0, -1]], [1e-5, 0.05], rtol=1e-6)
    np.testing.assert_allclose(np.diff(betas.astype(np.float64)), (0.05 - 1e-5) / 49, rtol=1e-4)
    exact = np.cumprod(1.0 - np.linspace(1e-5, 0.05, 50))
    np.testing.assert_allclose(ab, exact, rtol=2e-6)
    np.testing.assert_allclose(sig, np.sqrt(1.0 - ab.astype(np.float64)), rtol=1e-6)
    # 1 - alphas_bar in float32 (as the reference computes it) carries up to
    # ~0.6 % relative rounding at beta0 = 1e-5
    np.testing.assert_allclose(sig, np.sqrt(1.0 - exact), rtol=5e-3)
    np.testing.assert_allclose(sig[0], np.sqrt(1e-5), rtol=5e-3)
    assert np.all(np.diff(sig) > 0) and np.all(np.diff(ab) < 0)
    # the reverse step's identity alphas_bar[i] = alphas[i] alphas_bar[i - 1]
    np.testing.assert_allclose(ab[1:], alphas[1:] * ab[:-1], rtol=1e-6)


def test_jitter_offsets_ring():
    """J start offsets on a circle of the given radius, evenly spaced from the
    phase; one zero offset for J <= 1."""
    np.testing.assert_array_equal(pl.jitter_offsets(1, 0.05), np.zeros((1, 2)))
    np.testing.assert_array_equal(pl.jitter_offsets(0, 0.05), np.zeros((1, 2)))
    o = pl.jitter_offsets(4, 0.05)
    assert o.shape == (4, 2)
    np.testing.assert_allclose(np.hypot(o[:, 0], o[:, 1]), 0.05, rtol=1e-12)
    np.testing.assert_allclose(o.sum(axis=0), 0.0, atol=1e-15)
    np.testing.assert_allclose(o[0], [0.05, 0.0], atol=1e-15)
    np.testing.assert_allclose(o[1], [0.0, 0.05], atol=1e-15)
    r = pl.jitter_offsets(4, 0.05, np.pi / 4)
    np.testing.assert_allclose(r[0], [0.05 / np.sqrt(2), 0.05 / np.sqrt(2)], rtol=1e-12)
