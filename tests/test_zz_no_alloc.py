"""gcfg_dt_compute_noalloc (the data term with its scratch in a fixed device
array, for callers that cannot allocate) against gcfg_dt_compute on the same
images: every output plane bit for bit, on two calls in a row, the second on
a smaller image, so the second runs on scratch the first left behind."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(g, W, H, img, seeds, lam):
    d_rgb = g.DeviceBuffer.from_array(img)
    d_seeds = g.DeviceBuffer.from_array(seeds)
    outs = [g.DeviceBuffer(4 * W * H) for _ in range(6)]
    rc = g.lib().gcfg_dt_compute_noalloc(W, H, ctypes.c_void_p(d_rgb.ptr), ctypes.c_void_p(d_seeds.ptr), lam,
                                         *[ctypes.c_void_p(o.ptr) for o in outs], None)
    assert rc == 0, g.lib().gcfg_last_error()
    g.device_synchronize()
    return [o.download(np.int32, W * H).reshape(H, W) for o in outs]


def _reference(g, W, H, img, seeds, lam):
    dt = g.DataTerm(W, H)
    dt.set_lambda(lam)
    dt.set_image(img)
    dt.set_seeds(seeds)
    dt.compute()
    out = dt.links()
    dt.close()
    return out


def test_noalloc_equals_dataterm_twice(g):
    rng = np.random.default_rng(5)
    for W, H, lam, seed in ((1031, 777, 50, 1), (333, 201, 17, 2)):
        img = g.gen_ellipse_image(W, H, seed)
        seeds = g.gen_seed_map(W, H)
        seeds[rng.random((H, W)) < 0.01] = 2  # scattered background seeds as well
        print(f"[noalloc] {W}x{H} lambda {lam}", flush=True)
        got = _noalloc(g, W, H, img, seeds, lam)
        want = _reference(g, W, H, img, seeds, lam)
        for name, a, b in zip(("pos", "neg", "up", "down", "left", "right"), got, want):
            assert np.array_equal(a, b), (W, H, name)


def test_noalloc_rejects_oversize(g):
    rc = g.lib().gcfg_dt_compute_noalloc(8192, 8193, None, None, 50, None, None, None, None, None, None, None)
    assert rc == g.GCFG_ERR_INVALID
    assert b"2^26" in g.lib().gcfg_last_error()
