"""Device data term vs the oracle: histograms and beta exact, quantised
capacities exact, pre-quantisation values within 1e-5 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _device(g, W, H, img, seeds, lam=50):
    dt = g.DataTerm(W, H)
    if lam != 50:
        dt.set_lambda(lam)
    dt.set_image(img)
    dt.set_seeds(seeds)
    dt.compute()
    return dt


@pytest.mark.parametrize("shape,seed", [((320, 240), 1), ((97, 61), 2), ((1024, 768), 3), ((33, 500), 4)])
def test_ellipse_dataterm_exact(g, ora, shape, seed):
    W, H = shape
    img = g.gen_ellipse_image(W, H, seed)
    seeds = g.gen_seed_map(W, H)
    dt = _device(g, W, H, img, seeds)
    ref = ora.dataterm(W, H, img, seeds, 50)
    hist, counts = dt.histograms()
    assert np.array_equal(hist, ref["hist"]) and np.array_equal(counts, ref["counts"])
    bsum, pairs, beta = dt.beta()
    assert bsum == ref["beta_sum"] and pairs == ref["pairs"] and beta == ref["beta"]
    for name, a in zip(("pos", "neg", "up", "down", "left", "right"), dt.links()):
        assert np.array_equal(a, ref[name]), name
    tpos, tneg, rf, df = dt.unquantized()
    for a, b in ((tpos, ref["tpos_f"]), (tneg, ref["tneg_f"]), (rf, ref["right_f"]), (df, ref["down_f"])):
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        assert float(rel.max()) <= 1e-5
    dt.close()


def test_perlin_and_lambda(g, ora):
    W, H = 400, 300
    img = g.gen_perlin_image(W, H, 5)
    seeds = np.zeros((H, W), np.uint8)
    seeds[10:40, 10:40] = 1
    seeds[200:260, 300:380] = 2
    for lam in (50, 7, 1000):
        dt = _device(g, W, H, img, seeds, lam)
        ref = ora.dataterm(W, H, img, seeds, lam)
        for name, a in zip(("pos", "neg", "up", "down", "left", "right"), dt.links()):
            assert np.array_equal(a, ref[name]), (lam, name)
        dt.close()


def test_seed_rects_and_reset(g, ora):
    W, H = 128, 96
    img = g.gen_ellipse_image(W, H, 8)
    dt = g.DataTerm(W, H)
    dt.set_image(img)
    dt.add_seed_rect(54, 38, 20, 20, g.SEED_FG)
    dt.add_seed_rect(0, 0, 20, 20, g.SEED_BG)
    dt.add_seed_rect(W - 20, H - 20, 30, 30, g.SEED_BG)  # clipped at the image edge
    dt.compute()
    seeds = np.zeros((H, W), np.uint8)
    seeds[38:58, 54:74] = 1
    seeds[0:20, 0:20] = 2
    seeds[H - 20:, W - 20:] = 2
    ref = ora.dataterm(W, H, img, seeds, 50)
    hist, counts = dt.histograms()
    assert np.array_equal(hist, ref["hist"]) and list(counts) == [400, 800]
    dt.reset()
    dt.compute()
    _, counts = dt.histograms()
    assert counts.sum() == 0
    dt.close()


def test_constant_image(g):
    W, H = 50, 40
    img = np.full((H, W, 3), 90, np.uint8)
    dt = _device(g, W, H, img, np.zeros((H, W), np.uint8))
    pos, neg, up, down, left, right = dt.links()
    assert (right[:, :-1] == 50).all() and (right[:, -1] == 0).all()
    assert (up[1:] == 50).all() and (up[0] == 0).all()
    assert dt.beta()[2] == 0.0
    dt.close()
