"""Video loop on serving sessions (the per-frame loop of
ComplexApplication.cu:202-207): every frame's data term uses the colour model
of the session's first image, the residual graph is reparametrised onto the
new frame's capacities and solved warm, and each frame's flow and labels equal
a cold solve of the same capacities; edits between frames stay applied."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _drain(srv, n):
    out = {}
    while len(out) < n:
        for r in srv.poll(wait=True):
            out[r["request_id"]] = r
    return out


def _frames(g, W, H, seed, n):
    """A moving scene: the first image shifted by (3k, 2k) pixels, with fresh
    noise on a tenth of the pixels of every frame."""
    rng = np.random.default_rng(seed)
    base = g.gen_ellipse_image(W, H, seed)
    out = [base]
    for k in range(1, n):
        f = np.roll(base, (2 * k, 3 * k), axis=(0, 1)).copy()
        m = rng.random((H, W)) < 0.1
        f[m] = rng.integers(0, 256, (int(m.sum()), 3))
        out.append(f.astype(np.uint8))
    return out


def _cold(ora, frame, model, seeds, W, H, dpos, dneg):
    ref = ora.dataterm(W, H, frame, seeds, 50, model_rgb=model)
    E = (ref["up"], ref["down"], ref["left"], ref["right"])
    return ora.maxflow(W, H, ref["pos"] + dpos, ref["neg"] + dneg, edges=E)


@pytest.mark.parametrize("kernels,lazy,fmt", [(1, 0, 0), (1, 2, 0), (0, 0, 0), (1, 0, 1), (0, 0, 1)])
def test_video_frames_warm_equal_cold(g, ora, kernels, lazy, fmt):
    """fmt 1: labels packed one bit per pixel, which the scheduler writes to
    the request's pinned staging as each frame's solve finishes."""
    W, H = 320, 240
    frames = _frames(g, W, H, 3, 5)
    seeds = g.gen_seed_map(W, H)
    o = g.default_options()
    o.kernels = kernels
    o.lazy_caps = lazy
    srv = g.Server(1 << 30, o, label_format=fmt)
    rid0, sess, lab0 = srv.submit_image(frames[0], seeds)
    ids = [(rid0, lab0)]
    for f in frames[1:3]:
        ids.append(srv.submit_frame(sess, f))
    rid_e, lab_e = srv.submit_edit(sess, [[40, 30, 16, 16, 2], [150, 110, 16, 16, 1]], 1 << 16, shape=(H, W))
    for f in frames[3:]:
        ids.append(srv.submit_frame(sess, f))
    res = _drain(srv, len(ids) + 1)
    assert all(r["status"] == 0 for r in res.values())
    dpos = np.zeros((H, W), np.int64)
    dneg = np.zeros((H, W), np.int64)
    for k, (rid, label) in enumerate(ids):
        if k == 3:  # the edit was applied after frame 2
            dneg[30:46, 40:56] += 1 << 16
            dpos[110:126, 150:166] += 1 << 16
        rf, rl = _cold(ora, frames[k], frames[0], seeds, W, H, dpos, dneg)
        got = g.unpack_bits(label, W, H) if fmt == 1 else label
        assert res[rid]["flow"] == rf, k
        assert np.array_equal(got, rl), k
        assert 2000 < got.sum() < W * H // 2
    srv.close()


def test_video_sessions_in_the_mix(g, ora):
    """Two video sessions interleaved with one-off images: each frame's result
    equals the cold solve, whatever else is on the device."""
    shapes = [(256, 192), (400, 160)]
    seqs = [_frames(g, W, H, 20 + i, 4) for i, (W, H) in enumerate(shapes)]
    srv = g.Server(1 << 30)
    sess, ids = [], []
    for i, (W, H) in enumerate(shapes):
        rid, s, lab = srv.submit_image(seqs[i][0], g.gen_seed_map(W, H))
        sess.append(s)
        ids.append((i, 0, rid, lab))
    other = g.gen_ellipse_image(300, 300, 77)
    for k in range(1, 4):
        for i, (W, H) in enumerate(shapes):
            rid, lab = srv.submit_frame(sess[i], seqs[i][k])
            ids.append((i, k, rid, lab))
        srv.submit_image(other, g.gen_seed_map(300, 300))
    res = _drain(srv, len(ids) + 3)
    assert all(r["status"] == 0 for r in res.values())
    for i, k, rid, lab in ids:
        W, H = shapes[i]
        z = np.zeros((H, W), np.int64)
        rf, rl = _cold(ora, seqs[i][k], seqs[i][0], g.gen_seed_map(W, H), W, H, z, z)
        assert res[rid]["flow"] == rf and np.array_equal(lab, rl), (i, k)
    srv.close()


def test_frame_size_mismatch_is_an_error(g):
    W, H = 128, 96
    srv = g.Server(1 << 28)
    rid, sess, _ = srv.submit_image(g.gen_ellipse_image(W, H, 1), g.gen_seed_map(W, H))
    rid2, _ = srv.submit_frame(sess, g.gen_ellipse_image(W + 1, H, 2))
    res = _drain(srv, 2)
    assert res[rid]["status"] == 0 and res[rid2]["status"] == g.GCFG_ERR_INVALID
    srv.close()
