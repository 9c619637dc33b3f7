"""The C++ drop-in (include/GraphCut.h, include/DataTerm.h) driven the way the
reference's interactive application drives its own solver
(ComplexApplication.cu:105-135,180-207): one GC handle and one DataTerm reused
over frames, kernel-density seeds from clicks on the first frame, a constant
penalty (or the contrast n-links), labels on the host, selection drawn.
Results are checked against the oracle, frame by frame."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "tests", "drivers", "compat_app")


def _run(tmp_path, W, H, penalty, sigma, model, clicks, frames):
    assert os.path.exists(APP), "tests/drivers/compat_app is built by __graft_entry__.build()"
    src = tmp_path / "in.bin"
    dst = tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([W, H, penalty, sigma, model, len(clicks), len(frames)], np.int32).tofile(f)
        np.array(clicks, np.int32).reshape(-1).tofile(f)
        for fr in frames:
            np.ascontiguousarray(fr, np.uint8).tofile(f)
    print(f"[compat_app] {W}x{H} model {model} penalty {penalty} frames {len(frames)}", flush=True)
    subprocess.run([APP, str(src), str(dst)], check=True, timeout=120)
    raw = dst.read_bytes()
    n = W * H
    per = 8 + 3 * 4 * n + 3 * n
    assert len(raw) == per * len(frames)
    out = []
    for k in range(len(frames)):
        b = raw[k * per:(k + 1) * per]
        flow = int(np.frombuffer(b[:8], np.int64)[0])
        ints = np.frombuffer(b[8:8 + 12 * n], np.int32).reshape(3, H, W)
        shown = np.frombuffer(b[8 + 12 * n:], np.uint8).reshape(H, W, 3)
        out.append((flow, ints[0], ints[1], ints[2], shown))
    return out


@pytest.mark.parametrize("model", ["kde", "default"])
def test_kde_video_loop_penalty(g, ora, tmp_path, model):
    """The reference's colour model, asked for explicitly and as the drop-in's
    default (a driver that only switches headers gets the same t-links)."""
    W, H = 160, 120
    frames = [g.gen_ellipse_image(W, H, s) for s in (21, 22, 23)]
    # eight clicks inside the ellipse; a kernel width of 30 grey levels covers
    # the generator's colour noise (sigma 20 per channel)
    clicks = [(75, 55), (80, 60), (70, 58), (60, 50), (90, 65), (78, 40), (72, 75), (85, 52)]
    sigma, pen = 30, 3000
    res = _run(tmp_path, W, H, pen, sigma, g.DT_MODEL_KDE if model == "kde" else -1, clicks, frames)
    cols = np.concatenate([ora.kde_select_colours(frames[0], W, H, x, y) for (x, y) in clicks])
    for k, (flow, pos, neg, label, shown) in enumerate(res):
        ref = ora.dataterm_kde(W, H, frames[k], cols, sigma)
        assert np.array_equal(pos, ref["pos"]) and np.array_equal(neg, ref["neg"]), k
        assert (neg < 0).any()
        rflow, rlabel = ora.maxflow(W, H, pos, neg, penalty=pen)
        assert flow == rflow and np.array_equal(label, rlabel), k
        # most of the ellipse (area pi * 0.3W * 0.35H ~ 6330 px) is selected
        assert 4000 < label.sum() < 8000
        want = frames[k].copy()
        want[label == 1] = (0, 0, 255)  # K_DrawMap, DataTermKernels.cu:107-112
        assert np.array_equal(shown, want), k


def test_histogram_model_contrast_edges(g, ora, tmp_path):
    W, H = 130, 90
    frames = [g.gen_ellipse_image(W, H, 31)]
    clicks = [(60, 40), (62, 44)]
    (flow, pos, neg, label, _), = _run(tmp_path, W, H, 0, 5, g.DT_MODEL_HISTOGRAM, clicks, frames)
    seeds = np.zeros((H, W), np.uint8)
    for (x, y) in clicks:
        seeds[y:y + 5, x:x + 5] = 1
    ref = ora.dataterm(W, H, frames[0], seeds, 50)
    assert np.array_equal(pos, ref["pos"]) and np.array_equal(neg, ref["neg"])
    E = tuple(ref[k] for k in ("up", "down", "left", "right"))
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    assert flow == rflow and np.array_equal(label, rlabel)
