"""Pins the data-term oracle with closed-form special cases (CPU only).

The formulas are fixed in the oracle's header; these cases are ones where the
mathematics gives the answer directly, chosen so that plausible mistakes
(swapped pos/neg, missing +1 smoothing, wrong bin shift, wrong normaliser,
beta over the wrong pair set, transposed n-link directions) fail.
"""
import math

import numpy as np


def two_colour_image(W, H, a, b, mask):
    img = np.empty((H, W, 3), np.uint8)
    img[mask] = a
    img[~mask] = b
    return img


def test_histograms_and_tlinks_closed_form(ora):
    W, H = 40, 30
    fg, bg = (200, 80, 60), (60, 120, 180)
    mask = np.zeros((H, W), bool)
    mask[:, : W // 2] = True
    img = two_colour_image(W, H, fg, bg, mask)
    seeds = np.zeros((H, W), np.uint8)
    seeds[2:7, 2:9] = 1  # 35 foreground seeds in the fg half
    seeds[20:25, 30:36] = 2  # 30 background seeds in the bg half
    out = ora.dataterm(W, H, img, seeds, 50)
    nf, nb = 35, 30
    assert list(out["counts"]) == [nf, nb]
    hist = out["hist"]
    for c in range(3):
        assert hist[0, c, fg[c] >> 4] == nf and hist[0, c].sum() == nf
        assert hist[1, c, bg[c] >> 4] == nb and hist[1, c].sum() == nb
    # fg pixel: p_fg = ((nf+1)/(nf+16))^3, p_bg = (1/(nb+16))^3 (smoothing only)
    tneg_fg = -64 * 3 * math.log((nf + 1) / (nf + 16))
    tpos_fg = 64 * 3 * math.log(nb + 16)
    assert out["neg"][10, 5] == round(tneg_fg) and out["pos"][10, 5] == round(tpos_fg)
    assert abs(out["tneg_f"][10, 5] - tneg_fg) < 1e-9 and abs(out["tpos_f"][10, 5] - tpos_fg) < 1e-9
    # bg pixel: mirror image
    assert out["pos"][10, 35] == round(-64 * 3 * math.log((nb + 1) / (nb + 16)))
    assert out["neg"][10, 35] == round(64 * 3 * math.log(nf + 16))
    # strong preference where expected: fg pixels are expensive as background
    assert out["pos"][10, 5] > out["neg"][10, 5]


def test_model_image_histograms_frame_beta(ora):
    """A video session's later frame: the seed histograms come from the model
    image (the session's first frame), beta and the n-links from the frame.
    Frame and model are built so that every quantity differs between the two
    readings, and each expected value is a closed form."""
    W, H = 32, 8
    A, B, C = (200, 40, 40), (40, 200, 100), (120, 120, 120)  # bins (12,2,2), (2,12,6), (7,7,7)
    left = np.zeros((H, W), bool)
    left[:, :16] = True
    model = two_colour_image(W, H, A, B, left)  # model: A | B
    frame = two_colour_image(W, H, C, A, left)  # frame: C | A
    seeds = np.zeros((H, W), np.uint8)
    seeds[:, 2] = 1  # model colour A (the frame has C there)
    seeds[:, 20] = 2  # model colour B (the frame has A there)
    out = ora.dataterm(W, H, frame, seeds, 50, model_rgb=model)
    n = 8
    assert list(out["counts"]) == [n, n]
    hist = out["hist"]
    for c in range(3):
        assert hist[0, c, A[c] >> 4] == n and hist[0, c].sum() == n
        assert hist[1, c, B[c] >> 4] == n and hist[1, c].sum() == n
    # beta from the frame: only the 8 pairs across x = 15 | 16 differ, by |C - A|^2.
    pairs = (W - 1) * H + W * (H - 1)
    sq = sum((C[c] - A[c]) ** 2 for c in range(3))
    assert out["beta_sum"] == 8 * sq and out["pairs"] == pairs
    beta = 8 * sq / pairs
    # C pixels (left half of the frame): no seed colour shares a bin with C.
    assert out["pos"][3, 5] == round(64 * 3 * math.log(n + 16))
    assert out["neg"][3, 5] == round(64 * 3 * math.log(n + 16))
    # A pixels (right half): the foreground model's colour.
    assert out["neg"][3, 25] == round(-64 * 3 * math.log((n + 1) / (n + 16)))
    assert out["pos"][3, 25] == round(64 * 3 * math.log(n + 16))
    # n-links: lambda inside each half, the edge weight from the frame's beta.
    w_edge = round(50 * math.exp(-sq / (2 * beta)))
    assert out["right"][3, 15] == w_edge and out["left"][3, 16] == w_edge
    assert out["right"][3, 5] == 50 and out["down"][3, 25] == 50


def test_unseeded_class_is_uniform(ora):
    W, H = 12, 10
    img = np.full((H, W, 3), 77, np.uint8)
    seeds = np.zeros((H, W), np.uint8)
    out = ora.dataterm(W, H, img, seeds, 50)
    assert out["counts"].sum() == 0
    assert (out["pos"] == round(64 * 3 * math.log(16))).all()  # (1/16)^3
    assert (out["neg"] == out["pos"]).all()


def test_constant_image_beta_zero(ora):
    W, H = 9, 7
    img = np.full((H, W, 3), 130, np.uint8)
    out = ora.dataterm(W, H, img, np.zeros((H, W), np.uint8), 50)
    assert out["beta_sum"] == 0 and out["beta"] == 0.0
    assert (out["right"][:, :-1] == 50).all() and (out["right"][:, -1] == 0).all()
    assert (out["down"][:-1] == 50).all() and (out["down"][-1] == 0).all()
    assert (out["left"][:, 1:] == 50).all() and (out["left"][:, 0] == 0).all()
    assert (out["up"][1:] == 50).all() and (out["up"][0] == 0).all()


def test_checkerboard_beta_is_the_only_difference(ora):
    W, H = 11, 8
    a, b = (10, 200, 30), (40, 180, 90)
    s = sum((x - y) ** 2 for x, y in zip(a, b))
    yy, xx = np.mgrid[0:H, 0:W]
    img = two_colour_image(W, H, a, b, (xx + yy) % 2 == 0)
    out = ora.dataterm(W, H, img, np.zeros((H, W), np.uint8), 50)
    pairs = (W - 1) * H + W * (H - 1)
    assert out["pairs"] == pairs and out["beta_sum"] == s * pairs and out["beta"] == s
    w = round(50 * math.exp(-0.5))  # = 30
    assert w == 30
    assert (out["right"][:, :-1] == w).all() and (out["down"][:-1] == w).all()


def test_single_edge_image_directions(ora):
    """Left half one colour, right half another: only the column pair at the
    split is cut; directions and symmetry of the planes are checked."""
    W, H = 10, 6
    a, b = (0, 0, 0), (0, 30, 40)  # |a-b|^2 = 2500
    mask = np.zeros((H, W), bool)
    mask[:, :4] = True
    img = two_colour_image(W, H, a, b, mask)
    out = ora.dataterm(W, H, img, np.zeros((H, W), np.uint8), 50)
    pairs = (W - 1) * H + W * (H - 1)
    beta = 2500 * H / pairs
    assert abs(out["beta"] - beta) < 1e-12
    w_cut = round(50 * math.exp(-2500 / (2 * beta)))
    assert (out["right"][:, 3] == w_cut).all() and (out["left"][:, 4] == w_cut).all()
    assert (out["right"][:, :3] == 50).all() and (out["right"][:, 4:-1] == 50).all()
    assert (out["down"][:-1] == 50).all()
    assert np.array_equal(out["up"][1:], out["down"][:-1])
    assert np.array_equal(out["left"][:, 1:], out["right"][:, :-1])


def test_ellipse_image_consistency(ora, g):
    W, H = 320, 240
    img = g.gen_ellipse_image(W, H, 1)
    seeds = g.gen_seed_map(W, H)
    out = ora.dataterm(W, H, img, seeds, 50)
    assert list(out["counts"]) == [400, 1600]
    assert (out["hist"].sum(axis=2) == out["counts"][:, None]).all()
    # beta from its definition over all 4-neighbour pairs
    f = img.astype(np.int64)
    s = ((f[:, 1:] - f[:, :-1]) ** 2).sum() + ((f[1:] - f[:-1]) ** 2).sum()
    assert out["beta_sum"] == s
    # the ellipse interior prefers foreground, the corners background
    assert out["pos"][H // 2, W // 2] > out["neg"][H // 2, W // 2]
    assert out["neg"][5, 5] > out["pos"][5, 5]
    assert np.abs(out["tpos_f"] - out["pos"]).max() <= 0.5 + 1e-9
