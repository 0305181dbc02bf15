"""lbm_download_f_noalloc through the C ABI: the same [Q][N] values as
lbm_download_f (bit for bit), on two calls in a row on different lattices,
the second smaller (so nothing of the first call's staging survives into the
second's result), for each storage type and both distributions; lattices
beyond its static staging array are rejected with an error."""
import numpy as np
import pytest

import _pkg
from test_gpu_parity import random_state, say

pytestmark = pytest.mark.gpu


def lattice(L, shape, storage, seed):
    rng = np.random.default_rng(seed)
    dims = len(shape)
    mask = (rng.random(shape) < 0.3).astype(np.uint8)
    kw = dict(dims=dims, ni=shape[-1], nj=shape[-2], nk=shape[0] if dims == 3 else 1, nsubs=2,
              tau=(0.8, 0.75), gaccel=(1e-5, 0, 1e-5, 0, 0, 0), storage=storage,
              compute=0 if storage == 0 else 1)
    g = L.Lattice(mask=mask, **kw)
    f, c = random_state(rng, 19 if dims == 3 else 9, mask.size, 2)
    g.upload_f(0, f)
    g.upload_f(1, c)
    g.run(3)
    return g, mask


@pytest.mark.parametrize("storage", [0, 1, 2])
def test_noalloc_download_equals_download(storage):
    say(f"case no-alloc download storage={storage}")
    L = _pkg.load()
    for shape, seed in (((20, 16, 24), 1), ((6, 5, 7), 2), ((13, 11), 3)):
        g, mask = lattice(L, shape, storage, seed)
        for s in range(2):
            ref = g.download_f(s)
            got = g.download_f_noalloc(s)
            assert got.shape == ref.shape
            assert np.array_equal(got, ref), (shape, s)
            solid = mask.reshape(-1) == 1
            assert not got[:, solid].any()
        say(f"storage={storage} shape={shape}: equal")
        g.close()


def test_noalloc_download_rejects_large_lattice():
    L = _pkg.load()
    n = 77  # 19 * 77^3 = 8,674,127 doubles > 8,388,608
    g = L.Lattice(mask=np.zeros((n, n, n), np.uint8), dims=3, ni=n, nj=n, nk=n, tau=(0.8, 0.8))
    g.init_rest()
    with pytest.raises(L.LbmError, match="exceed the 8388608 of its static staging array"):
        g.download_f_noalloc(0)
    # the handle is unaffected
    assert np.array_equal(g.download_f(0)[:, :5], np.zeros((19, 5)))
    g.close()
