"""Pins the oracle's restatement of the reference's kernel-density colour
model (DataTermKernels.cu:19-31,61-83,116-150; DataTerm.cpp:97-137) on the CPU.

The oracle sums kernel values per pixel (a gather); the reference splats every
seed into a 256^3 float volume (a scatter) and reads the volume at the pixel's
colour. The checks below restate the scatter literally in numpy -- the
volume layout pdf[(c1*256 + c0)*256 + c2], float32 distances through sqrt,
the d < 101 support, float32 accumulation in seed order -- so a transposed
channel, a wrong support boundary, rounding instead of truncation or a lost
multiplicity fails; and the kernel values are checked against the normal
density they approximate.
"""
import math

import numpy as np
import pytest

N_PDF = 256


def test_kde_kernel_is_the_scaled_normal_density(ora):
    for sigma in (1, 3, 5, 8, 20):
        g = ora.kde_gauss(sigma).astype(np.float64)
        k = np.arange(101, dtype=np.float64)
        normal = np.exp(-0.5 * (k / sigma) ** 2) / (sigma * math.sqrt(2 * math.pi))
        # 2.507 stands in for sqrt(2 pi) = 2.50663 in the reference (DataTermKernels.cu:10)
        want = normal * math.sqrt(2 * math.pi) / 2.507
        # The reference rounds the exponent to float before expf, a relative
        # error of up to |exponent| * 2^-24 in the value.
        arg = 0.5 * (k / sigma) ** 2
        nz = want > 1e-30
        rel = np.abs(g[nz] - want[nz]) / want[nz]
        assert (rel <= 1e-6 + arg[nz] * 2.0 ** -23).all()
        assert g[0] == np.float32(1.0 / (sigma * 2.507))


def scatter_volume(seed_colours, gauss, lo, hi):
    """K_SelectAndGauss over the colour sub-cube [lo, hi]^3 (only those voxels
    are read below): for every seed, every voxel (x, y, z) with
    d = sqrtf(|s - (x, y, z)|^2) < 101 gets gauss[(int)d] added, in float32."""
    pdf = {}
    xs = np.arange(lo[0], hi[0] + 1)
    ys = np.arange(lo[1], hi[1] + 1)
    zs = np.arange(lo[2], hi[2] + 1)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    flat = ((Y * N_PDF + X) * N_PDF + Z).ravel()  # the reference's index, DataTermKernels.cu:147
    vol = np.zeros(flat.shape, np.float32)
    for s in seed_colours:
        d2 = ((float(s[0]) - X) ** 2 + (float(s[1]) - Y) ** 2 + (float(s[2]) - Z) ** 2).astype(np.float32)
        d = np.sqrt(d2).ravel()
        hit = d < 101
        vol[hit] = (vol[hit] + gauss[d[hit].astype(np.int64)]).astype(np.float32)
    for i, v in zip(flat, vol):
        pdf[int(i)] = v
    return pdf


def calc_terms(rgb, pdf):
    """K_CalcTerms: p = pdf[(y*N+x)*N+z] with (x, y, z) the pixel's bytes."""
    H, W, _ = rgb.shape
    pos = np.zeros((H, W), np.int64)
    neg = np.zeros((H, W), np.int64)
    p = np.zeros((H, W), np.float32)
    for yy in range(H):
        for xx in range(W):
            x, y, z = (int(v) for v in rgb[yy, xx])
            pv = np.float32(pdf.get((y * N_PDF + x) * N_PDF + z, 0.0))
            p[yy, xx] = pv
            pos[yy, xx] = math.trunc(300000.0 * float(pv))
            neg[yy, xx] = math.trunc(300000.0 * float(np.float32(1.0) - pv))
    return p, pos, neg


def test_kde_matches_reference_volume_scatter(ora):
    rng = np.random.default_rng(3)
    W, H = 24, 18
    lo, hi = (90, 60, 130), (150, 120, 190)
    img = np.stack([rng.integers(lo[c], hi[c] + 1, (H, W)) for c in range(3)], axis=-1).astype(np.uint8)
    # a near-uniform region, so that the clicks inside it push p above 1
    img[:12, :12] = (np.array([120, 90, 160]) + rng.integers(-2, 3, (12, 12, 3))).astype(np.uint8)
    # three clicks, one overlapping another (repeated pixels count twice), one at the edge
    cols = np.concatenate([ora.kde_select_colours(img, W, H, x, y) for (x, y) in ((3, 4), (5, 6), (21, 15))])
    assert cols.shape == (25 + 25 + 3 * 3, 3)
    sigma = 5
    g = ora.kde_gauss(sigma)
    pdf = scatter_volume(cols, g, lo, hi)
    p_ref, pos_ref, neg_ref = calc_terms(img, pdf)
    out = ora.dataterm_kde(W, H, img, cols, sigma)
    assert np.allclose(out["p"], p_ref, rtol=1e-5, atol=0)
    assert (out["p"] > 1).any() and (out["neg"] < 0).any()  # unnormalised density: neg < 0 occurs
    # truncated ints: the float volume's accumulation order may move a value
    # across an integer by at most the 1e-5 relative tolerance
    tol = np.ceil(300000.0 * 1e-5 * np.maximum(out["p"], 1.0)) + 1
    assert (np.abs(out["pos"] - pos_ref) <= tol).all()
    assert (np.abs(out["neg"] - neg_ref) <= tol).all()


def test_kde_single_seed_closed_form(ora):
    sigma = 5
    g = ora.kde_gauss(sigma)
    s = np.array([[100, 50, 200]], np.uint8)
    # pixel colours at known distances from the seed: 0, 3, 5 (3-4-0), sqrt(10200) (floor 100), 101, 102
    offs = [(0, 0, 0), (3, 0, 0), (0, 3, 4), (100, 10, 10), (101, 0, 0), (0, 0, -102)]
    assert 100 ** 2 + 10 ** 2 + 10 ** 2 == 10200
    img = np.array([[s[0].astype(int) + np.array(o) for o in offs]], np.int64)
    assert img.min() >= 0 and img.max() <= 255
    img = img.astype(np.uint8)
    out = ora.dataterm_kde(len(offs), 1, img, s, sigma)
    want_p = [g[0], g[3], g[5], g[100], 0.0, 0.0]
    assert np.array_equal(out["p"][0], np.array(want_p, np.float64))
    for i, pv in enumerate(want_p):
        pf = np.float32(pv)
        assert out["pos"][0, i] == math.trunc(300000.0 * float(pf))  # truncation, not rounding
        assert out["neg"][0, i] == math.trunc(300000.0 * float(np.float32(1.0) - pf))
    assert out["pos"][0, 4] == 0 and out["neg"][0, 4] == 300000  # outside the support


def test_kde_selection_order_and_clipping(ora):
    W, H = 7, 6
    img = np.arange(W * H * 3, dtype=np.uint8).reshape(H, W, 3)
    cols = ora.kde_select_colours(img, W, H, 4, 3)
    # x offset outer, y offset inner; x in 4..6 (7, 8 outside), y in 3..5 (6, 7 outside)
    want = [img[y, x] for x in range(4, 7) for y in range(3, 6)]
    assert np.array_equal(cols, np.array(want))
    assert ora.kde_select_colours(img, W, H, -10, -10).shape == (0, 3)


@pytest.mark.parametrize("seed", range(6))
def test_negative_tlinks_brute_force(ora, seed):
    """Negative t-links (the kernel-density term produces neg < 0): the oracle
    returns the minimum cut energy with the t-links as given and the union of
    the optimal source sets."""
    from conftest import random_instance
    from test_oracle import brute_force

    rng = np.random.default_rng(100 + seed)
    W, H = [(4, 4), (5, 4), (4, 6)][seed % 3]
    pos, neg, E = random_instance(rng, W, H, tmax=12, emax=8)
    pos = pos - rng.integers(0, 8, pos.shape).astype(np.int32)
    neg = neg - rng.integers(0, 14, neg.shape).astype(np.int32)
    assert (neg < 0).any() and (pos < 0).any()
    best, blab = brute_force(W, H, pos, neg, E)
    flow, label = ora.maxflow(W, H, pos, neg, edges=E)
    assert flow == best
    assert np.array_equal(label, blab)
