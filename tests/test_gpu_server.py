"""Serving: each result is the same alone and in the mix (bit-exact), warm-
started edits equal a cold solve of the edited instance, the HBM budget is
enforced, and edits wait for their session."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _req(g, W, H, seed):
    return g.gen_ellipse_image(W, H, seed), g.gen_seed_map(W, H)


def _drain(srv, n):
    out = {}
    while len(out) < n:
        for r in srv.poll(wait=True):
            out[r["request_id"]] = r
    return out


def _cold(g, ora, img, seeds, W, H, dpos=None, dneg=None):
    ref = ora.dataterm(W, H, img, seeds, 50)
    pos = ref["pos"] + (0 if dpos is None else dpos)
    neg = ref["neg"] + (0 if dneg is None else dneg)
    E = (ref["up"], ref["down"], ref["left"], ref["right"])
    return ora.maxflow(W, H, pos, neg, edges=E)


def test_alone_equals_mix_and_oracle(g, ora):
    shapes = [(300, 200), (128, 256), (511, 300), (256, 256), (90, 400)]
    reqs = [(W, H) + _req(g, W, H, 100 + i) for i, (W, H) in enumerate(shapes)]
    alone = []
    for W, H, img, seeds in reqs:
        srv = g.Server(1 << 30)
        rid, sess, label = srv.submit_image(img, seeds)
        r = _drain(srv, 1)[rid]
        assert r["status"] == 0
        alone.append((r["flow"], label.copy()))
        srv.close()
    srv = g.Server(1 << 30)
    ids = []
    for W, H, img, seeds in reqs:
        rid, sess, label = srv.submit_image(img, seeds)
        ids.append((rid, label))
    res = _drain(srv, len(reqs))
    for k, (rid, label) in enumerate(ids):
        assert res[rid]["status"] == 0
        assert res[rid]["flow"] == alone[k][0]
        assert np.array_equal(label, alone[k][1])
        W, H, img, seeds = reqs[k]
        rf, rl = _cold(g, ora, img, seeds, W, H)
        assert rf == alone[k][0] and np.array_equal(rl, alone[k][1])
    srv.close()


@pytest.mark.parametrize("lazy", [0, 2])
def test_warm_edit_equals_cold(g, ora, lazy):
    W, H = 360, 240
    img, seeds = _req(g, W, H, 7)
    o = g.default_options()
    o.lazy_caps = lazy  # 2: capacities written before each tile's first push pass
    srv = g.Server(1 << 30, o)
    rid, sess, label0 = srv.submit_image(img, seeds)
    _drain(srv, 1)
    rng = np.random.default_rng(3)
    dpos = np.zeros((H, W), np.int64)
    dneg = np.zeros((H, W), np.int64)
    for round_ in range(3):
        strokes = []
        for _ in range(int(rng.integers(1, 6))):
            x0, y0 = int(rng.integers(0, W - 16)), int(rng.integers(0, H - 16))
            cls = int(rng.integers(1, 3))
            strokes.append([x0, y0, 16, 16, cls])
            (dpos if cls == 1 else dneg)[y0:y0 + 16, x0:x0 + 16] += 1 << 16
        rid, label = srv.submit_edit(sess, strokes, 1 << 16, shape=(H, W))
        r = _drain(srv, 1)[rid]
        assert r["status"] == 0
        rf, rl = _cold(g, ora, img, seeds, W, H, dpos, dneg)
        assert r["flow"] == rf, round_
        assert np.array_equal(label, rl), round_
    srv.close()


def test_edits_queue_behind_their_image_and_mix(g, ora):
    W, H = 200, 150
    img, seeds = _req(g, W, H, 9)
    img2, seeds2 = _req(g, 256, 128, 10)
    srv = g.Server(1 << 30)
    rid0, sess, _ = srv.submit_image(img, seeds)
    strokes = [[50, 50, 16, 16, 2], [120, 70, 16, 16, 1]]
    rid1, label1 = srv.submit_edit(sess, strokes, 1 << 16, shape=(H, W))
    rid2, _, _ = srv.submit_image(img2, seeds2)
    rid3, label3 = srv.submit_edit(sess, [[10, 10, 16, 16, 1]], 1 << 16, shape=(H, W))
    res = _drain(srv, 4)
    assert all(r["status"] == 0 for r in res.values())
    dpos = np.zeros((H, W), np.int64)
    dneg = np.zeros((H, W), np.int64)
    dneg[50:66, 50:66] += 1 << 16
    dpos[70:86, 120:136] += 1 << 16
    rf, rl = _cold(g, ora, img, seeds, W, H, dpos, dneg)
    assert res[rid1]["flow"] == rf and np.array_equal(label1, rl)
    dpos[10:26, 10:26] += 1 << 16
    rf, rl = _cold(g, ora, img, seeds, W, H, dpos, dneg)
    assert res[rid3]["flow"] == rf and np.array_equal(label3, rl)
    srv.close()


def test_budget_refusal_and_eviction(g):
    W, H = 512, 512
    img, seeds = _req(g, W, H, 11)
    srv = g.Server(4 << 20)  # too small for one 512^2 request
    rid, sess, _ = srv.submit_image(img, seeds)
    r = _drain(srv, 1)[rid]
    assert r["status"] == g.GCFG_ERR_BUDGET
    srv.close()
    srv = g.Server(40 << 20)  # room for about one session at a time: idle ones get evicted
    ids = [srv.submit_image(*_req(g, W, H, 20 + i))[:2] for i in range(3)]
    res = _drain(srv, 3)
    assert all(res[rid]["status"] == 0 for rid, _ in ids)
    _, _, used = srv.load()
    assert used <= 40 << 20
    rid, _ = srv.submit_edit(ids[0][1], [[5, 5, 16, 16, 1]], 1 << 16, shape=(H, W))
    r = _drain(srv, 1)[rid]
    assert r["status"] == g.GCFG_ERR_STATE  # the first session was evicted
    srv.close()


def test_edit_budget(g):
    """A session takes edits while the sum of delta x strokes stays within
    2^29 (the int32 excess bound); the edit past it fails, the session stays
    usable."""
    W, H = 130, 70
    img, seeds = _req(g, W, H, 3)
    srv = g.Server(1 << 30)
    rid, sess, _ = srv.submit_image(img, seeds)
    assert _drain(srv, 1)[rid]["status"] == 0
    stroke = [[10, 10, 16, 16, 1]]
    rid, _ = srv.submit_edit(sess, stroke * 4, (1 << 27) - 1, shape=(H, W))
    assert _drain(srv, 1)[rid]["status"] == 0
    rid, _ = srv.submit_edit(sess, stroke, 8, shape=(H, W))
    assert _drain(srv, 1)[rid]["status"] == g.GCFG_ERR_INVALID
    rid, _ = srv.submit_edit(sess, stroke, 4, shape=(H, W))
    assert _drain(srv, 1)[rid]["status"] == 0
    srv.close()


def _strokes(rng, W, H, k):
    out = []
    for _ in range(k):
        out.append([int(rng.integers(0, W - 16)), int(rng.integers(0, H - 16)), 16, 16, int(rng.integers(1, 3))])
    return out


def _apply(dpos, dneg, strokes, delta):
    for x0, y0, w, h, cls in strokes:
        (dpos if cls == 1 else dneg)[y0:y0 + h, x0:x0 + w] += delta


@pytest.mark.parametrize("fmt", [0, 1])
def test_serving_sizes_vs_cold(g, ora, fmt):
    """Requests of 256^2..2048^2 (two of at least 1536^2) in one stream, with
    edits of 1-5 strokes of +2^16 on the two large sessions and on a small
    one, all submitted together: every result (flow and labels, int32 or
    packed) equals a cold solve of the same capacities by the oracle."""
    rng = np.random.default_rng(40 + fmt)
    shapes = [(2048, 2048), (1600, 1536), (1024, 1024), (300, 256), (700, 900), (256, 512)]
    srv = g.Server(8 << 30, label_format=fmt)
    reqs = []  # (rid, label, session index, strokes so far)
    imgs = []
    for i, (W, H) in enumerate(shapes):
        img, seeds = _req(g, W, H, 60 + i)
        imgs.append((W, H, img, seeds))
        rid, sess, label = srv.submit_image(img, seeds)
        reqs.append((rid, label, i, []))
        if i in (0, 1, 3):
            done = []
            for e in range(2):
                st = _strokes(rng, W, H, int(rng.integers(1, 6)))
                done = done + st
                rid, label = srv.submit_edit(sess, st, 1 << 16, shape=(H, W))
                reqs.append((rid, label, i, list(done)))
    print(f"[server] {len(reqs)} requests, label format {fmt}", flush=True)
    res = _drain(srv, len(reqs))
    for rid, label, i, strokes in reqs:
        W, H, img, seeds = imgs[i]
        assert res[rid]["status"] == 0, (i, res[rid])
        dpos = np.zeros((H, W), np.int64)
        dneg = np.zeros((H, W), np.int64)
        _apply(dpos, dneg, strokes, 1 << 16)
        rf, rl = _cold(g, ora, img, seeds, W, H, dpos, dneg)
        got = g.unpack_bits(label, W, H) if fmt == 1 else label
        assert res[rid]["flow"] == rf, (i, len(strokes))
        assert np.array_equal(got, rl), (i, len(strokes), int((got != rl).sum()))
    srv.close()


def test_pooled_buffers_serve_later_sessions(g, ora):
    """A closed session's buffers serve the next, smaller session: its result
    equals a cold solve (nothing of the earlier image leaks in)."""
    srv = g.Server(1 << 30)
    W, H = 400, 300
    img, seeds = _req(g, W, H, 71)
    rid, sess, _ = srv.submit_image(img, seeds)
    assert _drain(srv, 1)[rid]["status"] == 0
    srv.close_session(sess)
    for W2, H2, sd in ((390, 290, 72), (300, 380, 73)):
        img2, seeds2 = _req(g, W2, H2, sd)
        rid, sess2, label = srv.submit_image(img2, seeds2)
        r = _drain(srv, 1)[rid]
        assert r["status"] == 0
        rf, rl = _cold(g, ora, img2, seeds2, W2, H2)
        assert r["flow"] == rf and np.array_equal(label, rl)
        srv.close_session(sess2)
    srv.close()


def test_admission_while_chunks_run(g, ora):
    """New sessions admitted while chunks of solver steps for earlier ones
    are still queued on the device (non-blocking polls between
    submissions): each result equals a cold solve. This is the order in
    which a session's data term could see a seed map that work still queued
    ahead of it had not finished with."""
    shapes = [(2048, 1536), (640, 480), (512, 300), (900, 700), (333, 777), (1024, 1024), (256, 256), (800, 600)]
    srv = g.Server(8 << 30, label_format=g.LABELS_BITS)
    subs = []
    res = {}
    for i, (W, H) in enumerate(shapes):
        img, seeds = _req(g, W, H, 90 + i)
        rid, sess, label = srv.submit_image(img, seeds)
        subs.append((rid, label, W, H, img, seeds))
        for _ in range(3):  # admits, launches a chunk, returns without waiting
            for r in srv.poll(wait=False):
                res[r["request_id"]] = r
    while len(res) < len(subs):
        for r in srv.poll(wait=True):
            res[r["request_id"]] = r
    for rid, label, W, H, img, seeds in subs:
        assert res[rid]["status"] == 0
        rf, rl = _cold(g, ora, img, seeds, W, H)
        assert res[rid]["flow"] == rf, (W, H)
        assert np.array_equal(g.unpack_bits(label, W, H), rl), (W, H)
    srv.close()


@pytest.mark.parametrize("fmt", [0, 1])
def test_lent_labels(g, ora, fmt):
    """Requests submitted without a label buffer get their labels lent in the
    server's pinned memory (valid until the next poll); mixed in one stream
    with requests that bring their own buffers, and with edits of both kinds:
    every result equals a cold solve."""
    rng = np.random.default_rng(80 + fmt)
    shapes = [(1024, 768), (640, 480), (300, 300), (900, 1100), (256, 600)]
    srv = g.Server(4 << 30, label_format=fmt)
    reqs = []  # (rid, own label array or None, shape index, strokes so far)
    imgs = []
    for i, (W, H) in enumerate(shapes):
        img, seeds = _req(g, W, H, 110 + i)
        imgs.append((W, H, img, seeds))
        rid, sess, label = srv.submit_image(img, seeds, borrow=i % 2 == 0)
        assert (label is None) == (i % 2 == 0)
        reqs.append((rid, label, i, []))
        st = _strokes(rng, W, H, 3)
        rid, label = srv.submit_edit(sess, st, 1 << 16, shape=(H, W), borrow=i % 2 == 1)
        reqs.append((rid, label, i, st))
    res, lent = {}, {}
    while len(res) < len(reqs):
        for r in srv.poll(wait=True):
            res[r["request_id"]] = r
            if "labels" in r:
                lent[r["request_id"]] = np.array(r["labels"])  # a copy before the next poll
    for rid, label, i, strokes in reqs:
        W, H, img, seeds = imgs[i]
        assert res[rid]["status"] == 0, (i, res[rid])
        dpos = np.zeros((H, W), np.int64)
        dneg = np.zeros((H, W), np.int64)
        _apply(dpos, dneg, strokes, 1 << 16)
        rf, rl = _cold(g, ora, img, seeds, W, H, dpos, dneg)
        got = lent[rid] if label is None else label
        if fmt == 1:
            got = g.unpack_bits(got, W, H)
        assert res[rid]["flow"] == rf, (i, len(strokes))
        assert np.array_equal(got, rl), (i, len(strokes), int((got != rl).sum()))
    srv.close()


def test_label_formats_mixed_in_flight(g, ora):
    """Requests of both label formats in flight on one server: int32 images,
    then packed edits of those sessions and a packed new image, then int32
    edits again. A session's label destination follows each request's format
    (packed labels are written by the device into the request's staging,
    int32 labels copied after the solve), and every result equals a cold
    solve of the same capacities."""
    rng = np.random.default_rng(77)
    shapes = [(300, 200), (257, 130), (512, 384)]
    srv = g.Server(1 << 30)
    reqs = []  # (rid, label, format, session index, strokes so far)
    sess, imgs = [], []
    for i, (W, H) in enumerate(shapes):
        img, seeds = _req(g, W, H, 90 + i)
        imgs.append((W, H, img, seeds))
        rid, s, label = srv.submit_image(img, seeds)
        sess.append(s)
        reqs.append((rid, label, 0, i, []))
    done = [[] for _ in shapes]
    srv.set_label_format(g.LABELS_BITS)
    for i, (W, H) in enumerate(shapes):
        st = _strokes(rng, W, H, 2)
        done[i] = done[i] + st
        rid, label = srv.submit_edit(sess[i], st, 1 << 16, shape=(H, W))
        reqs.append((rid, label, 1, i, list(done[i])))
    W, H = 200, 160
    img, seeds = _req(g, W, H, 99)
    imgs.append((W, H, img, seeds))
    shapes.append((W, H))
    done.append([])
    rid, s, label = srv.submit_image(img, seeds)
    sess.append(s)
    reqs.append((rid, label, 1, 3, []))
    srv.set_label_format(g.LABELS_INT32)
    for i, (W, H) in enumerate(shapes):
        st = _strokes(rng, W, H, 1)
        done[i] = done[i] + st
        rid, label = srv.submit_edit(sess[i], st, 1 << 16, shape=(H, W))
        reqs.append((rid, label, 0, i, list(done[i])))
    print(f"[server] {len(reqs)} requests, formats mixed", flush=True)
    res = _drain(srv, len(reqs))
    for rid, label, fmt, i, strokes in reqs:
        W, H, img, seeds = imgs[i]
        assert res[rid]["status"] == 0, (i, fmt, res[rid])
        dpos = np.zeros((H, W), np.int64)
        dneg = np.zeros((H, W), np.int64)
        _apply(dpos, dneg, strokes, 1 << 16)
        rf, rl = _cold(g, ora, img, seeds, W, H, dpos, dneg)
        got = g.unpack_bits(label, W, H) if fmt == 1 else label
        assert res[rid]["flow"] == rf, (i, fmt, len(strokes))
        assert np.array_equal(got, rl), (i, fmt, len(strokes))
    srv.close()
