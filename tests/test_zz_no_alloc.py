"""The allocation-free connected-components entry point against the
allocating one, called through the C ABI: two calls in a row on one GPU,
the second with a smaller graph (the shared scratch keeps the first call's
contents), labels equal to the allocating entry point's and the oracle's."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import bonds as obonds

pytestmark = pytest.mark.gpu


def _graph(rng, n, nb):
    bi = rng.integers(0, n, nb).astype(np.int32)
    bj = rng.integers(0, n, nb).astype(np.int32)
    keep = bi != bj
    bi, bj = np.minimum(bi, bj)[keep], np.maximum(bi, bj)[keep]
    intact = (rng.random(len(bi)) < 0.7).astype(np.uint8)
    return bi, bj, intact


def _labels(fn, n, bi, bj, intact, n_iter):
    from icedem import _lib
    dev = torch.device("cuda", 0)
    tb = [torch.as_tensor(a).to(dev) for a in (bi, bj, intact)]
    out = torch.full((n,), -7, dtype=torch.int32, device=dev)
    strm = torch.cuda.current_stream().cuda_stream
    nb = len(bi)
    rc = fn(n, nb, tb[0].data_ptr() if nb else None, tb[1].data_ptr() if nb else None,
            tb[2].data_ptr() if nb else None, n_iter, out.data_ptr(), strm)
    _lib.check(rc, None, "connected components")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_equals_allocating_entry_point():
    from icedem import _lib
    lib = _lib.load()
    rng = np.random.default_rng(11)
    for n, nb, n_iter in ((50000, 60000, 64), (3000, 2500, 64), (700, 900, 5)):
        bi, bj, intact = _graph(rng, n, nb)
        a = _labels(lib.icedem_connected_components, n, bi, bj, intact, n_iter)
        b = _labels(lib.icedem_connected_components_noalloc, n, bi, bj, intact, n_iter)
        print(f"n={n} bonds={len(bi)} iters={n_iter}: components {len(np.unique(b))}", flush=True)
        np.testing.assert_array_equal(b, a)
        if n_iter == 64:
            np.testing.assert_array_equal(b, obonds.find_connected_components(n, bi, bj, intact.astype(bool)))


def test_noalloc_rejects_sizes_over_the_limit():
    from icedem import _lib
    lib = _lib.load()
    limit = 1 << 20
    rc = lib.icedem_connected_components_noalloc(limit + 1, 0, None, None, None, 64, None, None)
    assert rc == -3
    assert "exceeds the scratch limit" in lib.icedem_last_error(None).decode()
