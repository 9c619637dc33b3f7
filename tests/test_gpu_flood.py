"""Reachability rounds next to blocks that every capped round skips.

A capped relabel leaves a block at "every interior node below the cap" when
all its nodes reach a sink within the cap; later rounds skip such a block.
The reachability round that confirms a capped schedule has to treat those
nodes as reaching a sink even when they are not sinks themselves: a node
whose only route to a sink enters the skipped block through such a node
reaches the sink (label 0), and positive excess frozen at the cap behind it
still has to be pushed. These instances put the sinks of a region exactly
inside tiles (64 x 16) so that the region's tiles are skipped, with the rest
of the grid reaching them only through non-sink nodes, and compare flow and
labels with the oracle for several caps and relabel block shapes, and for
both capacity layouts (written by the build, or lazily by the push pass).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TW, TH = 64, 16


def _check(g, ora, W, H, pos, neg, penalty=0, edges=None, **opts):
    o = g.default_options()
    for k, v in opts.items():
        setattr(o, k, v)
    print(f"[flood] {W}x{H} {opts}", flush=True)
    flow, label, st = g.solve(W, H, pos, neg, edges=edges, penalty=penalty, options=o)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=edges, penalty=penalty)
    assert flow == rflow, (W, H, opts, flow, rflow)
    assert np.array_equal(label, rlabel), (W, H, opts, int((label != rlabel).sum()))
    return st


def half_plane(W, H, vertical, excess=0):
    """pos = neg = 0 up to column 64 (row 64 when vertical), sinks beyond;
    optionally positive excess at the far end, more than the cap away."""
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    if vertical:
        neg[65:, :] = 7
        if excess:
            pos[2:6, :] = excess
    else:
        neg[:, 65:] = 7
        if excess:
            pos[:, 2:6] = excess
    return pos, neg


OPTION_SETS = [dict(), dict(gr_cap=4), dict(gr_cap=16), dict(gr_cap=3), dict(gr_cap=31),
               dict(gr_cap=8, gr_block_w=2, gr_block_h=2), dict(gr_cap=8, gr_block_w=1, gr_block_h=4),
               dict(gr_cap=8, lazy_caps=1), dict(gr_cap=8, lazy_caps=2), dict(gr_cap=8, inner_iters=1)]


@pytest.mark.parametrize("opts", OPTION_SETS)
@pytest.mark.parametrize("excess", [0, 5])
def test_half_plane_horizontal(g, ora, opts, excess):
    W, H = 128, 16
    pos, neg = half_plane(W, H, False, excess)
    _check(g, ora, W, H, pos, neg, penalty=10, **opts)


@pytest.mark.parametrize("opts", OPTION_SETS)
@pytest.mark.parametrize("excess", [0, 5])
def test_half_plane_vertical(g, ora, opts, excess):
    W, H = 16, 128
    pos, neg = half_plane(W, H, True, excess)
    _check(g, ora, W, H, pos, neg, penalty=10, **opts)


@pytest.mark.parametrize("opts", OPTION_SETS[:6])
def test_boundary_inside_tile_row(g, ora, opts):
    """Sinks start one row below a tile edge (row 17), so tile row 1 is all
    below the cap while its first row holds no sink."""
    W, H = 3 * TW, 4 * TH
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    neg[TH + 1:, :] = 9
    pos[1:3, :] = 6
    _check(g, ora, W, H, pos, neg, penalty=4, **opts)


@pytest.mark.parametrize("opts", OPTION_SETS[:6])
def test_tile_aligned_rectangle(g, ora, opts):
    """A sink rectangle filling one tile whose outer ring holds no sink; the
    excess sits far from it, so every route passes through the ring."""
    W, H = 4 * TW, 4 * TH
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    x0, y0 = 2 * TW, TH
    neg[y0 + 1:y0 + TH - 1, x0 + 1:x0 + TW - 1] = 11
    pos[H - 4:H - 1, 1:8] = 9
    _check(g, ora, W, H, pos, neg, penalty=3, **opts)


@pytest.mark.parametrize("seed", range(6))
def test_random_tile_regions(g, ora, seed):
    """Random n-links; sink regions covering whole tiles minus their ring;
    random excess elsewhere."""
    rng = np.random.default_rng(seed)
    tx, ty = int(rng.integers(2, 5)), int(rng.integers(2, 6))
    W, H = tx * TW, ty * TH
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    for _ in range(int(rng.integers(1, 4))):
        a, b = int(rng.integers(0, tx)), int(rng.integers(0, ty))
        neg[b * TH + 1:(b + 1) * TH - 1, a * TW + 1:(a + 1) * TW - 1] = rng.integers(1, 30)
    free = neg == 0
    spots = rng.random((H, W)) < 0.02
    pos[free & spots] = rng.integers(1, 40, int((free & spots).sum()))
    E = tuple(rng.integers(1, 12, (H, W)).astype(np.int32) for _ in range(4))
    cap = [4, 8, 16][seed % 3]
    _check(g, ora, W, H, pos, neg, edges=E, gr_cap=cap)
