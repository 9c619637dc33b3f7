"""8-neighbourhood solver (NEIGHBORHOOD == 8, GraphCut.h:22,65-70) on the GPU
vs the exact oracle: bit-exact flow and "cannot reach the sink" labels, a
residual-graph certificate over all eight arcs, diagonal-only flow across tile
corners, and the GraphCut.h drop-in compiled with NEIGHBORHOOD 8."""
import os
import subprocess

import numpy as np
import pytest

from conftest import random_instance

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D8 = ((-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1))  # l r u d ul ur dl dr


def rand8(rng, W, H, tmax=40, emax=20):
    pos = rng.integers(0, tmax, (H, W)).astype(np.int32)
    neg = rng.integers(0, tmax, (H, W)).astype(np.int32)
    return pos, neg, tuple(rng.integers(0, emax, (H, W)).astype(np.int32) for _ in range(8))


def check8(g, ora, W, H, pos, neg, dirs=None, penalty=0, options=None):
    flow, label, st = g.solve8(W, H, pos, neg, dirs8=dirs, penalty=penalty, options=options)
    rflow, rlabel = ora.maxflow8(W, H, pos, neg, dirs8=dirs, penalty=penalty)
    assert flow == rflow, (W, H, flow, rflow)
    assert np.array_equal(label, rlabel), (W, H, int((label != rlabel).sum()))
    assert ora.cut_value8(W, H, pos, neg, label, dirs8=dirs, penalty=penalty) == flow
    return st


@pytest.mark.parametrize("shape", [(3, 3), (5, 4), (17, 9), (63, 15), (64, 16), (65, 17), (130, 33), (200, 70),
                                   (31, 120)])
def test_random8(g, ora, shape):
    W, H = shape
    rng = np.random.default_rng(7000 + W * 100 + H)
    for trial in range(3):
        pos, neg, dirs = rand8(rng, W, H)
        check8(g, ora, W, H, pos, neg, dirs)


@pytest.mark.parametrize("shape", [(17, 9), (65, 17), (200, 70)])
def test_random8_workgroup_kernels(g, ora, shape):
    """The workgroup-per-tile kernels (kernels=0) on the same instances."""
    W, H = shape
    rng = np.random.default_rng(7000 + W * 100 + H)
    o = g.default_options()
    o.kernels = 0
    for trial in range(2):
        pos, neg, dirs = rand8(rng, W, H)
        check8(g, ora, W, H, pos, neg, dirs, options=o)


def test_corner_flows8(g, ora):
    """Instances whose flow has to cross many tile corners diagonally:
    t-links only on a sparse lattice, n-links only on the diagonals, on a
    grid of 3 x 4 tiles (the corner nodes of every tile carry flow)."""
    rng = np.random.default_rng(88)
    W, H = 3 * 64 + 5, 4 * 16 + 3
    z = np.zeros((H, W), np.int32)
    for trial in range(3):
        pos = np.where(rng.random((H, W)) < 0.05, rng.integers(0, 500, (H, W)), 0).astype(np.int32)
        neg = np.where(rng.random((H, W)) < 0.05, rng.integers(0, 500, (H, W)), 0).astype(np.int32)
        diag = tuple(rng.integers(0, 80, (H, W)).astype(np.int32) for _ in range(4))
        check8(g, ora, W, H, pos, neg, (z, z, z, z) + diag)


def test_penalty8_and_glibc_sized(g, ora):
    rng = np.random.default_rng(3)
    W, H = 400, 400
    pos = rng.integers(0, 300, (H, W)).astype(np.int32)
    neg = rng.integers(0, 300, (H, W)).astype(np.int32)
    check8(g, ora, W, H, pos, neg, penalty=50)


def test_diagonal_chain_across_tile_corners(g, ora):
    """Flow only along down-right arcs, through the tile corners at (64, 16)
    and (128, 32): the 4-neighbourhood flow is 0, the 8-neighbourhood flow is
    the chain's bottleneck."""
    W, H = 160, 48
    z = np.zeros((H, W), np.int32)
    dr = z.copy()
    for k in range(50, 135):
        y = k - 49
        if y < H - 2 and k < W - 2:
            dr[y, k] = 30 + (k % 7)
    dr[20, 69] = 11  # bottleneck after the first corner
    pos, neg = z.copy(), z.copy()
    pos[1, 50] = 1000
    neg[H - 3, 50 + H - 4] = 1000
    dirs = (z, z, z, z, z, z, z, dr)
    flow, label, _ = g.solve8(W, H, pos, neg, dirs8=dirs)
    assert flow == 11
    check8(g, ora, W, H, pos, neg, dirs)
    assert ora.maxflow(W, H, pos, neg, edges=(z, z, z, z))[0] == 0


def test_heavy_tailed8(g, ora):
    W, H = 300, 200
    pos, neg = g.gen_stress_tlinks(W, H, 5)
    rng = np.random.default_rng(5)
    dirs = tuple(rng.integers(0, 50, (H, W)).astype(np.int32) for _ in range(8))
    check8(g, ora, W, H, pos, neg, dirs)


def test_negative_tlinks8(g, ora):
    rng = np.random.default_rng(8)
    W, H = 140, 70
    pos, neg, dirs = rand8(rng, W, H, tmax=200, emax=60)
    pos = pos - rng.integers(0, 120, pos.shape).astype(np.int32)
    neg = neg - rng.integers(0, 250, neg.shape).astype(np.int32)
    check8(g, ora, W, H, pos, neg, dirs)


def test_options8(g, ora):
    rng = np.random.default_rng(12)
    W, H = 150, 80
    pos, neg, dirs = rand8(rng, W, H, tmax=200, emax=60)
    # kernels=1 (the default): wavefront kernels, capped relabels with gr_cap;
    # kernels=0: the workgroup-per-tile kernels, always exact relabels.
    for kw in (dict(sweeps_per_gr=1, inner_iters=1), dict(sweeps_per_gr=8, inner_iters=64, use_graph=0),
               dict(gr_cap=4), dict(gr_cap=0), dict(gr_cap=16, inner_iters=1), dict(gr_cap=0, gr_rounds=1),
               dict(kernels=0), dict(kernels=0, sweeps_per_gr=2, inner_iters=4),
               dict(gr_block_w=2, gr_block_h=2)):  # relabel blocks are one tile in the 8-neighbourhood
        o = g.default_options()
        for k, v in kw.items():
            setattr(o, k, v)
        check8(g, ora, W, H, pos, neg, dirs, options=o)


def test_certificate8(g, ora):
    """Residual capacities over all eight arcs form a feasible preflow whose
    value is the cut of the labels, and the labels are the nodes with no
    residual path to a node with remaining sink capacity."""
    rng = np.random.default_rng(21)
    W, H = 170, 60
    pos, neg, dirs = rand8(rng, W, H, tmax=100, emax=40)
    gr = g.Graph(W, H, neighborhood=8)
    left, right, up, down, ul, ur, dl, dr = dirs
    gr.set_dataterms(pos, neg)
    gr.set_edges8(up, down, left, right, ul, ur, dl, dr)
    gr.build()
    label, st = gr.optimize()
    caps, exc = gr.read_state()
    caps8 = gr.read_diag_state()
    gr.close()
    res = np.concatenate([caps, caps8], axis=2).astype(np.int64)  # H x W x 8 in D8 order
    inter = np.zeros((H, W), bool)
    inter[1:-1, 1:-1] = True
    orig = np.stack(dirs, axis=2).astype(np.int64)
    for k, (dx, dy) in enumerate(D8):
        ys, xs = np.mgrid[0:H, 0:W]
        ok = inter & (xs + dx > 0) & (xs + dx < W - 1) & (ys + dy > 0) & (ys + dy < H - 1)
        orig[..., k] = np.where(ok, orig[..., k], 0)
    assert (res >= 0).all()
    # capacity conservation per arc pair: r(u,v) + r(v,u) == c(u,v) + c(v,u)
    for k, (dx, dy) in enumerate(D8):
        o = (1, 0, 3, 2, 7, 6, 5, 4)[k]
        for y in range(1, H - 1):
            x = np.arange(1, W - 1)
            xx, yy = x + dx, y + dy
            m = (xx > 0) & (xx < W - 1) & (yy > 0) & (yy < H - 1)
            x, xx = x[m], xx[m]
            assert np.array_equal(res[y, x, k] + res[yy, xx, o], orig[y, x, k] + orig[yy, xx, o])
    # excess: pos - neg + net inflow, summing to the flow identity
    assert int(exc[inter].sum()) == int(pos[inter].astype(np.int64).sum() - neg[inter].astype(np.int64).sum())
    assert st.flow == int(pos[inter].astype(np.int64).sum()) - int(np.maximum(exc[inter], 0).sum())
    # labels == cannot reach a node with excess < 0 through residual arcs
    reach = inter & (exc < 0)
    frontier = reach.copy()
    while frontier.any():
        new = np.zeros_like(reach)
        for k, (dx, dy) in enumerate(D8):
            # u -> v residual in direction k with v reached: u = v - (dx, dy)
            ys, xs = np.nonzero(frontier)
            us, ux = ys - dy, xs - dx
            m = (us > 0) & (us < H - 1) & (ux > 0) & (ux < W - 1)
            us, ux = us[m], ux[m]
            ok = res[us, ux, k] > 0
            new[us[ok], ux[ok]] = True
        new &= ~reach
        reach |= new
        frontier = new
    assert np.array_equal(label.astype(bool), inter & ~reach)
    assert ora.cut_value8(W, H, pos, neg, label, dirs8=dirs) == st.flow


def test_input_range_limits8(g, ora):
    """8-neighbourhood limits: n-links in [0, 2^27), |pos - neg| <= 2^29."""
    rng = np.random.default_rng(78)
    W, H = 66, 34
    big = 1 << 29
    pos = np.where(rng.random((H, W)) < 0.5, big, 0).astype(np.int32)
    neg = np.where(pos == 0, big, 0).astype(np.int32)
    dirs = tuple(np.where(rng.random((H, W)) < 0.7, (1 << 27) - 1, rng.integers(0, 1000, (H, W))).astype(np.int32)
                 for _ in range(8))
    check8(g, ora, W, H, pos, neg, dirs)
    d2 = tuple(d.copy() for d in dirs)
    d2[6][9, 9] = 1 << 27
    with pytest.raises(g.GcfgError):
        g.solve8(W, H, pos, neg, dirs8=d2)


def test_n8_api_errors(g):
    W, H = 20, 20
    z = np.zeros((H, W), np.int32)
    gr = g.Graph(W, H)  # 4-neighbourhood
    with pytest.raises(g.GcfgError):
        gr.set_edges8(*([z] * 8))
    gr.close()
    gr = g.Graph(W, H, neighborhood=8)
    gr.set_dataterms(z, z)
    gr.set_edges(z, z, z, z)  # side planes only: the diagonals are missing
    with pytest.raises(g.GcfgError):
        gr.build()
    gr.close()


def test_compat_driver_neighborhood8(g, ora, tmp_path):
    """GraphCut.h compiled with NEIGHBORHOOD 8: GC_Init / GC_SetPenalty / GC_SetGraph /
    GC_Optimize on kernel-density t-links, against the 8-neighbourhood oracle."""
    app = os.path.join(ROOT, "tests", "drivers", "compat_app8")
    assert os.path.exists(app), "tests/drivers/compat_app8 is built by __graft_entry__.build()"
    W, H = 150, 110
    frame = g.gen_ellipse_image(W, H, 41)
    clicks = [(70, 50), (75, 55), (65, 53), (80, 45), (72, 62)]
    sigma, pen = 30, 2000
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([W, H, pen, sigma, g.DT_MODEL_KDE, len(clicks), 1], np.int32).tofile(f)
        np.array(clicks, np.int32).tofile(f)
        frame.tofile(f)
    print(f"[compat_app8] {W}x{H} penalty {pen}", flush=True)
    subprocess.run([app, str(src), str(dst)], check=True, timeout=120)
    raw = dst.read_bytes()
    n = W * H
    flow = int(np.frombuffer(raw[:8], np.int64)[0])
    pos, neg, label = np.frombuffer(raw[8:8 + 12 * n], np.int32).reshape(3, H, W)
    cols = np.concatenate([ora.kde_select_colours(frame, W, H, x, y) for (x, y) in clicks])
    ref = ora.dataterm_kde(W, H, frame, cols, sigma)
    assert np.array_equal(pos, ref["pos"]) and np.array_equal(neg, ref["neg"])
    rflow, rlabel = ora.maxflow8(W, H, pos, neg, penalty=pen)
    assert flow == rflow and np.array_equal(label, rlabel)
    assert label.sum() > 1000
    f4, _ = ora.maxflow(W, H, pos, neg, penalty=pen)
    assert f4 != flow  # the diagonal arcs change the cut


@pytest.mark.parametrize("cfg", [dict(), dict(gr_cap=2, gr_cap_late=31), dict(inner_iters=1)])
def test_heavy_tailed8_partial_rounds(g, ora, cfg):
    """An 8-connected instance that needs more than the four capped rounds:
    growing capped rounds, partial exact rounds (whose blocks queue themselves
    for a sweep that reads their neighbours' distances) and a full exact round
    at the end give the exact flow and labels."""
    W, H = 480, 360
    pos, neg = g.gen_stress_tlinks(W, H, 17)
    rng = np.random.default_rng(17)
    dirs = tuple(rng.integers(0, 40, (H, W)).astype(np.int32) for _ in range(8))
    o = g.default_options()
    for k, v in cfg.items():
        setattr(o, k, v)
    st = check8(g, ora, W, H, pos, neg, dirs, options=o)
    if not cfg:
        assert st.gr_rounds > 5, st.gr_rounds
