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


# ---------------------------------------------------------------------------
# The reference's kernel-density model (DataTermKernels.cu:19-150).

def _kde_image(g, W, H, seed):
    img = g.gen_ellipse_image(W, H, seed)
    return img


def _kde_device(g, W, H, img, clicks, sigma=5):
    dt = g.DataTerm(W, H)
    dt.set_model(g.DT_MODEL_KDE)
    if sigma != 5:
        dt.set_sigma(sigma)
    dt.set_image(img)
    for (x, y) in clicks:
        dt.kde_select(x, y)
    dt.compute()
    return dt


def _clicks(W, H, many):
    c = [(W // 2, H // 2), (W // 2 + 3, H // 2 + 2), (W // 3, H // 2), (W - 3, H - 2), (-2, 1)]
    if many:  # more clicks inside the ellipse: the unnormalised density exceeds 1 there
        c += [(W // 2 - 8, H // 2 - 6), (W // 2 + 10, H // 2 + 5), (W // 2 - 4, H // 2 + 9), (W // 2 + 6, H // 2 - 9)]
    return c


@pytest.mark.parametrize("shape,sigma,many", [((320, 240), 5, False), ((97, 61), 3, False),
                                              ((1024, 768), 8, False), ((320, 240), 30, True),
                                              ((640, 480), 24, True)])
def test_kde_dataterm_exact(g, ora, shape, sigma, many):
    W, H = shape
    img = _kde_image(g, W, H, 11)
    clicks = _clicks(W, H, many)
    dt = _kde_device(g, W, H, img, clicks, sigma)
    cols = np.concatenate([ora.kde_select_colours(img, W, H, x, y) for (x, y) in clicks])
    assert dt.kde_count() == len(cols)
    ref = ora.dataterm_kde(W, H, img, cols, sigma)
    pos, neg = dt.links()[:2]
    assert np.array_equal(pos, ref["pos"]) and np.array_equal(neg, ref["neg"])
    if many:
        assert (neg < 0).any()  # p > 1 inside the clicked region, as in the reference
    tpos, tneg, _, _ = dt.unquantized()
    for a, b in ((tpos, ref["tpos_f"]), (tneg, ref["tneg_f"])):
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        assert float(rel.max()) <= 1e-5
    dt.close()


def test_kde_colours_captured_at_selection(g, ora):
    """selectOnMap reads the image at click time (DataTermKernels.cu:133-136):
    a later setImage (the next video frame) keeps the density."""
    W, H = 160, 120
    a = _kde_image(g, W, H, 1)
    b = _kde_image(g, W, H, 2)
    dt = g.DataTerm(W, H)
    dt.set_model(g.DT_MODEL_KDE)
    dt.set_image(a)
    dt.kde_select(70, 50)
    dt.kde_select(10, 10)
    dt.set_image(b)
    dt.compute()
    cols = np.concatenate([ora.kde_select_colours(a, W, H, 70, 50), ora.kde_select_colours(a, W, H, 10, 10)])
    ref = ora.dataterm_kde(W, H, b, cols, 5)
    pos, neg = dt.links()[:2]
    assert np.array_equal(pos, ref["pos"]) and np.array_equal(neg, ref["neg"])
    dt.reset()
    assert dt.kde_count() == 0
    dt.compute()
    pos, neg = dt.links()[:2]
    assert (pos == 0).all() and (neg == 300000).all()
    dt.close()


def test_kde_segmentation_negative_tlinks(g, ora):
    """Kernel-density t-links (neg < 0 where p > 1) with a constant penalty,
    as ComplexApplication.cu:127-135 uses them: exact flow and labels."""
    W, H = 200, 150
    img = _kde_image(g, W, H, 5)
    clicks = [(95, 70), (100, 75), (90, 72), (80, 60), (110, 80), (98, 55), (92, 88), (105, 66)]
    dt = _kde_device(g, W, H, img, clicks, sigma=30)
    pos, neg = dt.links()[:2]
    assert (neg < 0).any()
    for pen in (0, 700, 7000):
        flow, label, _ = g.solve(W, H, pos, neg, penalty=pen)
        rflow, rlabel = ora.maxflow(W, H, pos, neg, penalty=pen)
        assert flow == rflow and np.array_equal(label, rlabel), pen
    dt.close()


def _odd_shapes():
    # widths that are not multiples of 4 (rows start at every byte alignment),
    # widths below 4 (several rows per group of four pixels), one-pixel images,
    # sizes that end inside a chunk of 4096 pixels and exactly at one.
    return [(1, 1), (2, 3), (3, 7), (5, 2), (7, 9), (13, 5), (64, 64), (65, 63), (97, 61), (333, 17), (1023, 5),
            (4096, 1), (1, 4097)]


@pytest.mark.parametrize("lam", [50, 3])
def test_batch_dataterm_exact(g, ora, lam):
    """One launch per pass over images of every alignment: equal to the
    single-image path and to the oracle, bit for bit."""
    rng = np.random.default_rng(lam)
    images = []
    for i, (W, H) in enumerate(_odd_shapes()):
        img = g.gen_ellipse_image(W, H, 40 + i) if W >= 8 and H >= 8 else \
            rng.integers(0, 256, (H, W, 3)).astype(np.uint8)
        seeds = (rng.integers(0, 3, (H, W)) * (rng.random((H, W)) < 0.3)).astype(np.uint8)
        images.append((img, seeds if i % 3 else None))
    b = g.DataTermBatch(images, lam)
    for rep in range(2):  # the per-image counters are cleared for the next compute
        b.compute()
        for i, ((img, seeds), (W, H)) in enumerate(zip(images, b.shapes)):
            sd = seeds if seeds is not None else np.zeros((H, W), np.uint8)
            ref = ora.dataterm(W, H, img, sd, lam)
            got = b.links(i)
            for name, a in zip(("pos", "neg", "up", "down", "left", "right"), got):
                assert np.array_equal(a, ref[name]), (rep, i, W, H, name)
            if rep == 0:
                dt = _device(g, W, H, img, sd, lam)
                for a, c in zip(dt.links(), got):
                    assert np.array_equal(a, c), (i, W, H)
                dt.close()
    b.close()


@pytest.mark.parametrize("seed", range(4))
def test_every_colour_bin(g, ora, seed):
    """Images holding all 16^3 colour bins, with random seed histograms: every
    t-link table entry the device builds is compared with the oracle's
    per-pixel f64 evaluation."""
    rng = np.random.default_rng(100 + seed)
    W, H = 128, 64  # 8192 pixels: each bin twice, at random offsets inside the bin
    idx = np.arange(W * H) % 4096
    img = np.stack([((idx >> 8) & 15) * 16, ((idx >> 4) & 15) * 16, (idx & 15) * 16], -1).astype(np.int32)
    img = (img + rng.integers(0, 16, img.shape)).astype(np.uint8).reshape(H, W, 3)
    img = img.reshape(-1, 3)[rng.permutation(W * H)].reshape(H, W, 3)
    seeds = np.zeros((H, W), np.uint8)
    k = rng.integers(1, 3000)
    pick = rng.choice(W * H, size=int(k), replace=False)
    seeds.reshape(-1)[pick] = rng.integers(1, 3, int(k))
    dt = _device(g, W, H, img, seeds)
    ref = ora.dataterm(W, H, img, seeds, 50)
    for name, a in zip(("pos", "neg", "up", "down", "left", "right"), dt.links()):
        assert np.array_equal(a, ref[name]), name
    hist, counts = dt.histograms()
    assert np.array_equal(hist, ref["hist"]) and np.array_equal(counts, ref["counts"])
    dt.close()
