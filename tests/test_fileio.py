"""Host-side file formats of the reference's drivers (no GPU): the Middlebury
2-label MRF text file (solveMRF.cu:10-65, multiplyMRF.cpp:64-92), the
data-term file and the tiled constant-penalty MRF built from it
(multiplyMRF.cpp:94-145, Application.cu:15-24), and the PPM/PGM label images
(solveMRF.cu:113-127, Application.cu:55-75)."""
import numpy as np
import pytest


def _expected_planes(pos, neg, hcue, vcue, mult):
    """The arrays a reader following solveMRF.cu must produce from a file
    written by multiplyMRF.cpp's writer, computed independently in numpy."""
    H, W = pos.shape
    WW = W * mult
    cols = np.arange(WW)
    P = pos[:, cols % W]
    N = neg[:, cols % W]
    hc = np.zeros((H, WW), np.int32)
    hc[:, :WW - 1] = hcue[:, cols[:WW - 1] % (W - 1)]
    vc = np.zeros((H, WW), np.int32)
    vc[:H - 1, :] = vcue[:H - 1, cols % W]
    hc[H - 1, :] = 0  # solveMRF.cu:60-63 zeroes the last row of both cues
    vc[H - 1, :] = 0
    return WW, H, P, N, vc, vc, hc, hc


@pytest.mark.parametrize("mult", [1, 3])
def test_mrf_roundtrip(g, tmp_path, mult):
    rng = np.random.default_rng(mult)
    H, W = 5, 7
    pos, neg, hcue, vcue = (rng.integers(0, 1000, (H, W)).astype(np.int32) for _ in range(4))
    path = str(tmp_path / "m.txt")
    g.mrf_write(path, pos, neg, hcue, vcue, mult=mult)
    got = g.mrf_load(path)
    want = _expected_planes(pos, neg, hcue, vcue, mult)
    assert got[:2] == want[:2]
    for a, b in zip(got[2:], want[2:]):
        assert np.array_equal(a, b)


def test_mrf_known_file(g, tmp_path):
    """A 3x3 file written by hand: header, 9 ground-truth values, pos, neg,
    2 horizontal cues per row, 3 vertical cues per row but the last."""
    text = ("3 3 2\n" + "0 " * 9 + "\n" + "1 2 3 4 5 6 7 8 9\n" + "9 8 7 6 5 4 3 2 1\n"
            + "10 11  12 13  14 15\n" + "20 21 22  23 24 25\n")
    path = tmp_path / "k.txt"
    path.write_text(text)
    W, H, pos, neg, up, down, left, right = g.mrf_load(str(path))
    assert (W, H) == (3, 3)
    assert pos.ravel().tolist() == [1, 2, 3, 4, 5, 6, 7, 8, 9]
    assert neg.ravel().tolist() == [9, 8, 7, 6, 5, 4, 3, 2, 1]
    assert left.tolist() == [[10, 11, 0], [12, 13, 0], [0, 0, 0]] and np.array_equal(left, right)
    assert up.tolist() == [[20, 21, 22], [23, 24, 25], [0, 0, 0]] and np.array_equal(up, down)


def test_mrf_rejects_multilabel_and_truncation(g, tmp_path):
    p = tmp_path / "b.txt"
    p.write_text("2 2 3 " + "0 " * 40)
    with pytest.raises(g.GcfgError):
        g.mrf_load(str(p))
    p.write_text("3 3 2 " + "0 " * 20)
    with pytest.raises(g.GcfgError):
        g.mrf_load(str(p))


def test_label_ppm_text_format(g, tmp_path):
    lab = np.array([[1, 0], [0, 1]], np.int32)
    p = tmp_path / "l.ppm"
    g.write_label_ppm(str(p), lab)
    assert p.read_text() == "P3 2 2 255 \n255 255 255 0 0 0 \n0 0 0 255 255 255 \n"
    rgb = np.array([[[1, 2, 3], [4, 5, 6]], [[7, 8, 9], [10, 11, 12]]], np.uint8)
    g.write_label_ppm(str(p), lab, rgb)
    assert p.read_text() == "P3 2 2 255 \n1 2 3 0 0 0 \n0 0 0 10 11 12 \n"
    # and it reads back as an image
    back = g.read_ppm(str(p))
    assert back.shape == (2, 2, 3) and back[1, 1].tolist() == [10, 11, 12] and back[0, 1].tolist() == [0, 0, 0]


def test_ppm_p6_and_p3_read(g, tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (4, 5, 3)).astype(np.uint8)
    p6 = tmp_path / "a.ppm"
    p6.write_bytes(b"P6\n# comment\n5 4\n255\n" + img.tobytes())
    assert np.array_equal(g.read_ppm(str(p6)), img)
    p3 = tmp_path / "b.ppm"
    p3.write_text("P3 5 4 255\n" + " ".join(str(v) for v in img.ravel()) + "\n")
    assert np.array_equal(g.read_ppm(str(p3)), img)


def test_label_pgm(g, tmp_path):
    lab = np.array([[0, 1, 1], [1, 0, 0]], np.int32)
    p = tmp_path / "l.pgm"
    g.write_label_pgm(str(p), lab)
    assert p.read_bytes() == b"P5\n3 2\n255\n" + bytes([0, 255, 255, 255, 0, 0])


def test_dat_known_file_and_roundtrip(g, tmp_path):
    """'height width', then every neg value, then every pos value."""
    p = tmp_path / "face.dat"
    p.write_text("2 3\n1 2 3\n4 5 6\n10 20 30 40 50 60\n")
    W, H, pos, neg = g.dat_load(str(p))
    assert (W, H) == (3, 2)
    assert neg.tolist() == [[1, 2, 3], [4, 5, 6]] and pos.tolist() == [[10, 20, 30], [40, 50, 60]]
    rng = np.random.default_rng(5)
    pos2, neg2 = rng.integers(-50, 900, (2, 9, 13)).astype(np.int32)
    q = str(tmp_path / "r.dat")
    g.dat_write(q, pos2, neg2)
    W, H, a, b = g.dat_load(q)
    assert (W, H) == (13, 9) and np.array_equal(a, pos2) and np.array_equal(b, neg2)
    p.write_text("2 3\n1 2 3\n4 5 6\n10 20\n")
    with pytest.raises(g.GcfgError):
        g.dat_load(str(p))


@pytest.mark.parametrize("out", [(7, 5), (20, 11), (4, 3)])
def test_tiled_penalty_mrf(g, tmp_path, out):
    """writeRihannaMRFData's layout read back by the solveMRF reader: data
    terms repeat the source (column x mod W, row y mod H), every cue equals
    the penalty, and the reader's zeroed last row and last column apply."""
    rng = np.random.default_rng(out[0])
    H, W = 5, 7
    pos, neg = rng.integers(0, 1000, (2, H, W)).astype(np.int32)
    path = str(tmp_path / "t.txt")
    g.mrf_write_tiled(path, out[0], out[1], pos, neg, penalty=600)
    OW, OH, P, N, up, down, left, right = g.mrf_load(path)
    assert (OW, OH) == out
    ys, xs = np.meshgrid(np.arange(OH), np.arange(OW), indexing="ij")
    assert np.array_equal(P, pos[ys % H, xs % W]) and np.array_equal(N, neg[ys % H, xs % W])
    hc = np.full((OH, OW), 600, np.int32)
    hc[:, OW - 1] = 0
    hc[OH - 1, :] = 0
    vc = np.full((OH, OW), 600, np.int32)
    vc[OH - 1, :] = 0
    assert np.array_equal(left, hc) and np.array_equal(right, hc)
    assert np.array_equal(up, vc) and np.array_equal(down, vc)
    text = open(path).read().split()
    assert text[:3] == [str(out[0]), str(out[1]), "2"] and text[3:3 + OW * OH] == ["1"] * (OW * OH)
