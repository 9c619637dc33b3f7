"""Plain drop-in solver (GC_* path) on the GPU vs the exact oracle.

Bar: bit-exact flow value and bit-exact label map (the "cannot reach the
sink" set, unique over all maximum flows) on identical capacities.
"""
import numpy as np
import pytest

from conftest import debug_kat, glibc_rand_instance, random_instance

pytestmark = pytest.mark.gpu


def _check(g, ora, W, H, pos, neg, edges=None, penalty=0, options=None):
    flow, label, st = g.solve(W, H, pos, neg, edges=edges, penalty=penalty, options=options)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=edges, penalty=penalty)
    assert flow == rflow, (W, H, flow, rflow)
    assert np.array_equal(label, rlabel), (W, H, int((label != rlabel).sum()))
    assert ora.cut_value(W, H, pos, neg, label, edges=edges, penalty=penalty) == flow
    return st


def test_debug_kat(g, ora):
    W, H, pos, neg = debug_kat()
    flow, label, _ = g.solve(W, H, pos, neg, penalty=50)
    assert flow == 7
    assert label.sum() == 0


def _opts(g, kernels, **kw):
    o = g.default_options()
    o.kernels = kernels
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@pytest.mark.parametrize("kernels", [0, 1])
@pytest.mark.parametrize("shape", [(3, 3), (5, 4), (17, 9), (63, 15), (64, 16), (65, 17), (130, 33), (200, 70),
                                   (31, 120)])
def test_random_small(g, ora, shape, kernels):
    W, H = shape
    rng = np.random.default_rng(W * 1000 + H)
    for trial in range(3):
        pos, neg, E = random_instance(rng, W, H)
        _check(g, ora, W, H, pos, neg, edges=E, options=_opts(g, kernels))


def test_degenerate_shapes(g, ora):
    for W, H in ((1, 1), (2, 7), (9, 2), (3, 1)):
        pos = np.full((H, W), 5, np.int32)
        neg = np.zeros((H, W), np.int32)
        flow, label, _ = g.solve(W, H, pos, neg, penalty=3)
        assert flow == 0 and label.sum() == 0


def test_penalty_mode(g, ora):
    rng = np.random.default_rng(3)
    W, H = 150, 90
    pos, neg, _ = random_instance(rng, W, H, tmax=300, edges=False)
    _check(g, ora, W, H, pos, neg, penalty=50)


@pytest.mark.parametrize("kernels", [0, 1])
def test_debug_random_400(g, ora, kernels):
    """Debug.cu:60-74 random instance (glibc rand, seed 1), penalty 50."""
    pos, neg = glibc_rand_instance(1)
    _check(g, ora, 400, 400, pos, neg, penalty=50, options=_opts(g, kernels))


def serpentine(W, H, cap=9):
    """Every row is a corridor; row r connects to row r+1 only at one column,
    alternating between the left and right ends, so the only source-sink path
    snakes through the whole grid (length ~ W * H)."""
    up = np.zeros((H, W), np.int32)
    down = np.zeros((H, W), np.int32)
    left = np.full((H, W), cap, np.int32)
    right = np.full((H, W), cap, np.int32)
    for r in range(1, H - 2):
        gap = W - 2 if r % 2 else 1
        down[r, gap] = cap
        up[r + 1, gap] = cap
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    pos[1, 1:4] = 40
    neg[H - 2, 1:4] = 40
    return pos, neg, (up, down, left, right)


@pytest.mark.parametrize("kernels", [0, 1])
def test_long_paths(g, ora, kernels):
    """Serpentine corridor: long BFS distances across many tiles."""
    W, H = 100, 24
    pos, neg, (up, down, left, right) = serpentine(W, H)
    assert ora.maxflow(W, H, pos, neg, edges=(up, down, left, right))[0] == 9
    _check(g, ora, W, H, pos, neg, edges=(up, down, left, right), options=_opts(g, kernels))


@pytest.mark.parametrize("kernels", [0, 1])
def test_heavy_tailed(g, ora, kernels):
    W, H = 300, 200
    pos, neg = g.gen_stress_tlinks(W, H, 11)
    img = g.gen_perlin_image(W, H, 12)
    dt = g.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(np.zeros((H, W), np.uint8))
    dt.compute()
    _, _, up, down, left, right = dt.links()
    _check(g, ora, W, H, pos, neg, edges=(up, down, left, right), options=_opts(g, kernels))


def test_ellipse_c0(g, ora):
    """320x240 ellipse image through the device data term, then the plain solver."""
    W, H = 320, 240
    img = g.gen_ellipse_image(W, H, 1)
    seeds = g.gen_seed_map(W, H)
    dt = g.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(seeds)
    dt.compute()
    pos, neg, up, down, left, right = dt.links()
    st = _check(g, ora, W, H, pos, neg, edges=(up, down, left, right))
    assert st.steps > 0


@pytest.mark.parametrize("opts", [dict(sweeps_per_gr=1, inner_iters=1), dict(sweeps_per_gr=2, inner_iters=8),
                                  dict(sweeps_per_gr=32, inner_iters=256, steps_per_poll=3, use_graph=0),
                                  dict(sweeps_per_gr=4, inner_iters=64, kernels=0),
                                  dict(sweeps_per_gr=16, inner_iters=200, kernels=1, steps_per_poll=5)])
def test_options_do_not_change_result(g, ora, opts):
    rng = np.random.default_rng(99)
    W, H = 140, 75
    pos, neg, E = random_instance(rng, W, H, tmax=200, emax=60)
    o = g.default_options()
    for k, v in opts.items():
        setattr(o, k, v)
    _check(g, ora, W, H, pos, neg, edges=E, options=o)


def test_negative_nlink_rejected(g):
    W, H = 20, 20
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    E = tuple(np.ones((H, W), np.int32) for _ in range(4))
    E[2][5, 5] = -3
    with pytest.raises(g.GcfgError):
        g.solve(W, H, pos, neg, edges=E)
    with pytest.raises(g.GcfgError):  # |pos - neg| beyond 2^29 would leave int32 excess no headroom
        g.solve(W, H, np.full((H, W), 1 << 29, np.int32), np.full((H, W), -8, np.int32), penalty=1)


@pytest.mark.parametrize("kernels", [1, 0])
def test_input_range_limits(g, ora, kernels):
    """At the documented limits (|pos - neg| = 2^29, n-links 2^28 - 1) the
    excess of a node can approach 2^31 during the solve; results stay exact.
    One step past either limit is rejected."""
    o = g.default_options()
    o.kernels = kernels
    rng = np.random.default_rng(77)
    W, H = 70, 40
    big = 1 << 29
    pos = np.where(rng.random((H, W)) < 0.5, big, 0).astype(np.int32)
    neg = np.where(pos == 0, big, 0).astype(np.int32)
    E = tuple(np.where(rng.random((H, W)) < 0.7, (1 << 28) - 1, rng.integers(0, 1000, (H, W))).astype(np.int32)
              for _ in range(4))
    flow, label, _ = g.solve(W, H, pos, neg, edges=E, options=o)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    assert flow == rflow and np.array_equal(label, rlabel)
    with pytest.raises(g.GcfgError):
        p2 = pos.copy()
        p2[5, 5], neg[5, 5] = big, -1
        g.solve(W, H, p2, neg, edges=E, options=o)
    neg[5, 5] = 0
    E2 = tuple(e.copy() for e in E)
    E2[1][7, 9] = 1 << 28
    with pytest.raises(g.GcfgError):
        g.solve(W, H, pos, neg, edges=E2, options=o)
    with pytest.raises(g.GcfgError):
        g.solve(W, H, pos, neg, penalty=1 << 28, options=o)


@pytest.mark.parametrize("kernels", [1, 0])
def test_negative_tlinks_accepted(g, ora, kernels):
    """t-links of either sign, as the reference takes them (excess = pos - neg,
    GraphCutKernels.cu:196-197): the flow is the minimum cut energy with the
    t-links as given."""
    rng = np.random.default_rng(31)
    W, H = 150, 90
    pos, neg, E = random_instance(rng, W, H, tmax=200, emax=60)
    pos = pos - rng.integers(0, 120, pos.shape).astype(np.int32)
    neg = neg - rng.integers(0, 250, neg.shape).astype(np.int32)
    o = g.default_options()
    o.kernels = kernels
    _check(g, ora, W, H, pos, neg, edges=E, options=o)
    _check(g, ora, W, H, pos, neg, penalty=40, options=o)


def test_optimize_before_build_is_state_error(g):
    gr = g.Graph(10, 10)
    with pytest.raises(g.GcfgError) as e:
        gr.optimize()
    assert e.value.rc == g.GCFG_ERR_STATE
    gr.close()


def test_residual_certificate(g, ora):
    """The returned residual state is a feasible preflow whose value equals the
    cut value of the returned labels, and the labels are exactly the nodes that
    cannot reach a node with remaining sink capacity."""
    rng = np.random.default_rng(5)
    W, H = 190, 130
    pos, neg, E = random_instance(rng, W, H, tmax=100, emax=40)
    gr = g.Graph(W, H)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    label, st = gr.optimize()
    caps, exc = gr.read_state()
    gr.close()
    _verify_certificate(W, H, pos, neg, E, caps, exc, label, st.flow, ora)


def _verify_certificate(W, H, pos, neg, E, caps, exc, label, flow, ora):
    up, down, left, right = E
    inter = np.zeros((H, W), bool)
    inter[1:-1, 1:-1] = True
    # original capacities with the border rule
    c = np.zeros((H, W, 4), np.int64)
    c[..., 0], c[..., 1], c[..., 2], c[..., 3] = left, right, up, down
    c[~inter] = 0
    c[:, 1, 0] = 0
    c[:, W - 2, 1] = 0
    c[1, :, 2] = 0
    c[H - 2, :, 3] = 0
    r = caps.astype(np.int64)
    assert (r >= 0).all()
    # antisymmetric flow: r(u->v) + r(v->u) == c(u->v) + c(v->u)
    assert np.array_equal(r[:, 1:, 0] + r[:, :-1, 1], c[:, 1:, 0] + c[:, :-1, 1])
    assert np.array_equal(r[1:, :, 2] + r[:-1, :, 3], c[1:, :, 2] + c[:-1, :, 3])
    # conservation: e = pos - neg - net outflow
    net = (c - r).sum(axis=2)
    e_expect = np.where(inter, pos.astype(np.int64) - neg - net, 0)
    assert np.array_equal(exc.astype(np.int64), e_expect)
    assert flow == int(pos[inter].sum()) - int(np.maximum(exc[inter], 0).sum())
    assert ora.cut_value(W, H, pos, neg, label, edges=E) == flow
    # labels == cannot reach a node with e < 0 through residual arcs
    reach = inter & (exc < 0)
    while True:
        nxt = reach.copy()
        nxt[:, 1:] |= (r[:, 1:, 0] > 0) & reach[:, :-1]
        nxt[:, :-1] |= (r[:, :-1, 1] > 0) & reach[:, 1:]
        nxt[1:, :] |= (r[1:, :, 2] > 0) & reach[:-1, :]
        nxt[:-1, :] |= (r[:-1, :, 3] > 0) & reach[1:, :]
        nxt &= inter
        if np.array_equal(nxt, reach):
            break
        reach = nxt
    assert np.array_equal(label.astype(bool), inter & ~reach)


def test_rebuild_and_label_target_change(g, ora):
    """Building twice, then solving into a caller device buffer (a new label
    destination) gives the same exact result and terminates."""
    rng = np.random.default_rng(21)
    W, H = 120, 66
    pos, neg, E = random_instance(rng, W, H, tmax=90, emax=30)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    gr = g.Graph(W, H)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    gr.build()
    d_label = g.DeviceBuffer(4 * W * H)
    st = gr.optimize_device(d_label)
    assert st.flow == rflow
    assert np.array_equal(d_label.download(np.int32, W * H).reshape(H, W), rlabel)
    gr.build()
    label, st2 = gr.optimize()
    assert st2.flow == rflow and np.array_equal(label, rlabel)
    gr.close()


def test_profile_mode_reports_kernel_time(g, ora):
    rng = np.random.default_rng(4)
    W, H = 200, 100
    pos, neg, E = random_instance(rng, W, H, tmax=90, emax=30)
    o = g.default_options()
    o.profile = 1
    flow, label, st = g.solve(W, H, pos, neg, edges=E, options=o)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    assert flow == rflow and np.array_equal(label, rlabel)
    assert st.pr_kernel_ms > 0 and st.gr_kernel_ms > 0 and st.launches >= st.steps
    assert st.build_kernel_ms > 0
    assert st.pr_iters == 0 and st.gr_iters == 0  # counters only with profile 2
    o.profile = 2
    flow2, label2, st2 = g.solve(W, H, pos, neg, edges=E, options=o)
    assert flow2 == rflow and np.array_equal(label2, rlabel)
    assert st2.pr_iters > 0 and st2.gr_iters > 0 and st2.gr_kernel_ms > 0


GR_BLOCKS = [(1, 1), (1, 2), (1, 4), (1, 8), (1, 16), (2, 2), (2, 4), (2, 8), (4, 4)]


@pytest.mark.parametrize("block", GR_BLOCKS)
def test_gr_block_shapes(g, ora, block):
    """Every global-relabel block shape (BFS exchanged through LDS inside a
    block) gives the exact result, on shapes with partial blocks, a corridor
    that snakes through many tiles and heavy-tailed capacities."""
    o = g.default_options()
    o.gr_block_w, o.gr_block_h = block
    rng = np.random.default_rng(block[0] * 100 + block[1])
    for W, H in ((200, 70), (130, 33), (65, 300), (5, 4)):
        pos, neg, E = random_instance(rng, W, H, tmax=150, emax=50)
        _check(g, ora, W, H, pos, neg, edges=E, options=o)
    W, H = 100, 40
    pos, neg, E = serpentine(W, H)
    _check(g, ora, W, H, pos, neg, edges=E, options=o)
    W, H = 260, 180
    pos, neg = g.gen_stress_tlinks(W, H, 5)
    _, _, E = random_instance(rng, W, H, emax=80)
    _check(g, ora, W, H, pos, neg, edges=E, options=o)


def test_unsupported_gr_block_rejected(g):
    gr = g.Graph(10, 10)
    o = g.default_options()
    o.gr_block_w, o.gr_block_h = 3, 3
    with pytest.raises(g.GcfgError):
        gr.set_options(o)
    gr.close()


def _bfs_to_sinks(W, H, caps, exc):
    """Exact BFS distance of every interior node to a node with excess < 0
    along residual arcs (numpy level-synchronous relaxation)."""
    inter = np.zeros((H, W), bool)
    inter[1:-1, 1:-1] = True
    INF = 0x3FFFFFFF
    d = np.where(inter & (exc < 0), 0, INF).astype(np.int64)
    r = caps.astype(np.int64)
    while True:
        nd = d.copy()
        nd[:, 1:] = np.minimum(nd[:, 1:], np.where(r[:, 1:, 0] > 0, d[:, :-1] + 1, INF))
        nd[:, :-1] = np.minimum(nd[:, :-1], np.where(r[:, :-1, 1] > 0, d[:, 1:] + 1, INF))
        nd[1:, :] = np.minimum(nd[1:, :], np.where(r[1:, :, 2] > 0, d[:-1, :] + 1, INF))
        nd[:-1, :] = np.minimum(nd[:-1, :], np.where(r[:-1, :, 3] > 0, d[1:, :] + 1, INF))
        nd = np.where(inter, np.minimum(nd, INF), INF)
        if np.array_equal(nd, d):
            return d
        d = nd


@pytest.mark.parametrize("cfg", [dict(kernels=0), dict(kernels=1), dict(kernels=1, gr_block_w=1, gr_block_h=4),
                                 dict(kernels=1, gr_block_w=2, gr_block_h=8)])
def test_final_distances_are_exact_bfs(g, ora, cfg):
    """The solve ends with a converged global relabel: the stored distance of
    every interior node equals its BFS distance to the sink nodes in the final
    residual graph (checks the tile relaxation and its row scans directly)."""
    o = g.default_options()
    o.gr_cap = 0  # every relabel exact
    for k, v in cfg.items():
        setattr(o, k, v)
    rng = np.random.default_rng(17)
    cases = [(230, 150) + random_instance(rng, 230, 150, tmax=120, emax=40)]
    W, H = 100, 40
    cases.append((W, H) + serpentine(W, H))
    for W, H, pos, neg, E in cases:
        gr = g.Graph(W, H, o)
        gr.set_dataterms(pos, neg)
        gr.set_edges(*E)
        gr.build()
        label, st = gr.optimize()
        caps, exc = gr.read_state()
        dist = gr.read_distances()
        gr.close()
        inter = np.zeros((H, W), bool)
        inter[1:-1, 1:-1] = True
        want = _bfs_to_sinks(W, H, caps, exc)
        assert np.array_equal(dist[inter].astype(np.int64), want[inter]), (W, H, int((dist[inter] != want[inter]).sum()))
        assert (want[inter] < 0x3FFFFFFF).any() and (want[inter] > 1).any()


@pytest.mark.parametrize("cap", [1, 4, 16, 64])
def test_capped_relabels(g, ora, cap):
    """Capped intermediate global relabels (BFS fronts stopped at `cap`,
    nodes beyond frozen), confirmed by a reachability flood (followed by an
    exact round where an excess node can still reach a sink), still end with
    the exact flow and labels: long corridors, random and heavy-tailed
    instances, and the final reachability equals a BFS on the residual graph."""
    o = g.default_options()
    o.gr_cap = cap
    rng = np.random.default_rng(cap)
    for W, H in ((200, 70), (65, 300), (5, 4)):
        pos, neg, E = random_instance(rng, W, H, tmax=150, emax=50)
        _check(g, ora, W, H, pos, neg, edges=E, options=o)
    W, H = 100, 40
    pos, neg, E = serpentine(W, H)
    _check(g, ora, W, H, pos, neg, edges=E, options=o)
    W, H = 260, 180
    pos, neg = g.gen_stress_tlinks(W, H, 5)
    _, _, E = random_instance(rng, W, H, emax=80)
    _check(g, ora, W, H, pos, neg, edges=E, options=o)
    o.gr_block_w, o.gr_block_h = 2, 4
    gr = g.Graph(W, H, o)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    label, st = gr.optimize()
    caps, exc = gr.read_state()
    dist = gr.read_distances()
    gr.close()
    inter = np.zeros((H, W), bool)
    inter[1:-1, 1:-1] = True
    # A schedule finished within its capped rounds ends with a reachability
    # flood: distance 0 for the nodes that can reach a sink node, 0x3fffffff
    # for the others (the labels). One that outlasted them (relabels exact
    # from the fifth round on) ends with exact BFS distances.
    want = _bfs_to_sinks(W, H, caps, exc)[inter]
    reach = want < 0x3FFFFFFF
    assert np.array_equal(dist[inter] < 0x3FFFFFFF, reach)
    if not (dist[inter][reach] == 0).all():
        assert st.gr_rounds > 4
        assert np.array_equal(dist[inter].astype(np.int64), want)
    assert st.flow == ora.maxflow(W, H, pos, neg, edges=E)[0]


@pytest.mark.parametrize("cfg", [dict(), dict(gr_cap=8, gr_cap_late=8), dict(gr_cap=6, gr_cap_late=2),
                                 dict(gr_cap=2, gr_cap_late=31), dict(inner_iters=1),
                                 dict(gr_block_w=2, gr_block_h=2), dict(gr_cap=0)])
def test_cap_schedule_and_partial_rounds(g, ora, cfg):
    """The first relabel capped at gr_cap, the next three at gr_cap_late
    (a larger late cap recomputes every live block; a smaller one is raised
    to gr_cap), then exact rounds that recompute only the blocks near the
    last pushes, and a full exact round at the end: a heavy-tailed instance
    that needs more than five rounds gives the exact flow and labels, and its
    final distances are an exact BFS of the returned residual graph."""
    W, H = 640, 480
    pos, neg = g.gen_stress_tlinks(W, H, 21)
    img = g.gen_perlin_image(W, H, 22)
    dt = g.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(np.zeros((H, W), np.uint8))
    dt.compute()
    _, _, up, down, left, right = dt.links()
    E = (up, down, left, right)
    o = _opts(g, 1, **cfg)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    gr = g.Graph(W, H, o)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    label, st = gr.optimize()
    caps, exc = gr.read_state()
    dist = gr.read_distances()
    gr.close()
    assert st.flow == rflow
    assert np.array_equal(label, rlabel), int((label != rlabel).sum())
    if not cfg:
        assert st.gr_rounds > 5, st.gr_rounds  # partial exact rounds ran
    inter = np.zeros((H, W), bool)
    inter[1:-1, 1:-1] = True
    want = _bfs_to_sinks(W, H, caps, exc)[inter]
    reach = want < 0x3FFFFFFF
    assert np.array_equal(dist[inter] < 0x3FFFFFFF, reach)
    if not (dist[inter][reach] == 0).all():  # ended with an exact round, not a reachability flood
        assert np.array_equal(dist[inter].astype(np.int64), want)


def test_caller_stream_ordering(g, ora):
    """With a caller stream set, a device copy of the labels queued on that
    stream right after optimize_device sees the final labels; results equal
    the oracle."""
    rng = np.random.default_rng(33)
    W, H = 300, 120
    pos, neg, E = random_instance(rng, W, H, tmax=100, emax=30)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    s = g.Stream()
    gr = g.Graph(W, H)
    gr.set_stream(s.ptr)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    d_label = g.DeviceBuffer(4 * W * H)
    d_copy = g.DeviceBuffer(4 * W * H)
    st = gr.optimize_device(d_label)
    s.copy_async(d_copy.ptr, d_label.ptr, 4 * W * H)
    s.synchronize()
    assert st.flow == rflow
    assert np.array_equal(d_copy.download(np.int32, W * H).reshape(H, W), rlabel)
    gr.set_stream(0)
    gr.build()
    label, st2 = gr.optimize()
    assert st2.flow == rflow and np.array_equal(label, rlabel)
    gr.close()
    s.close()


def test_middlebury_file_solve(g, ora, tmp_path):
    """solveMRF.cu's workflow: a 2-label Middlebury file (written here by the
    multiplyMRF.cpp layout, tiled twice) is loaded with the reference's cue
    mapping and solved by the drop-in path; the result equals the oracle on
    the loaded planes, and the label image is written as solveMRF.cu does."""
    rng = np.random.default_rng(44)
    H, W = 60, 70
    pos, neg, hcue, vcue = (rng.integers(0, 200, (H, W)).astype(np.int32) for _ in range(4))
    path = str(tmp_path / "m.txt")
    g.mrf_write(path, pos, neg, hcue, vcue, mult=2)
    W2, H2, p, n, up, down, left, right = g.mrf_load(path)
    assert (W2, H2) == (2 * W, H)
    st = _check(g, ora, W2, H2, p, n, edges=(up, down, left, right))
    flow, label, _ = g.solve(W2, H2, p, n, edges=(up, down, left, right))
    out = tmp_path / "labelMRF.ppm"
    g.write_label_ppm(str(out), label)
    head = out.read_text().split("\n", 1)[0]
    assert head == f"P3 {W2} {H2} 255 "
    assert st.steps > 0


def test_face_dat_and_tiled_mrf_solve(g, ora, tmp_path):
    """Application.cu's workflow (a face.dat file of data terms solved with
    penalty 600) and multiplyMRF.cpp's Rihanna path (the same data terms
    tiled into a larger constant-penalty MRF file, read back by the
    solveMRF.cu reader): both equal the oracle."""
    rng = np.random.default_rng(45)
    H, W = 220, 150
    yy, xx = np.mgrid[0:H, 0:W]
    inside = ((xx - W / 2) / (0.3 * W)) ** 2 + ((yy - H / 2) / (0.35 * H)) ** 2 < 1
    base = rng.integers(0, 2000, (H, W))
    pos = np.where(inside, base + 3000, base).astype(np.int32)
    neg = np.where(inside, base, base + 3000).astype(np.int32)
    dat = str(tmp_path / "face.dat")
    g.dat_write(dat, pos, neg)
    W1, H1, p1, n1 = g.dat_load(dat)
    assert (W1, H1) == (W, H)
    st = _check(g, ora, W, H, p1, n1, penalty=600)
    mrf = str(tmp_path / "tiled.txt")
    g.mrf_write_tiled(mrf, 3 * W, H + 37, p1, n1, penalty=600)
    W2, H2, p2, n2, up, down, left, right = g.mrf_load(mrf)
    assert (W2, H2) == (3 * W, H + 37)
    st2 = _check(g, ora, W2, H2, p2, n2, edges=(up, down, left, right))
    assert st.steps > 0 and st2.steps > 0


@pytest.mark.parametrize("inner", [1, 4])
@pytest.mark.parametrize("lazy", [1, 2])
@pytest.mark.parametrize("shape", [(5, 4), (130, 33), (257, 140)])
def test_lazy_capacities(g, ora, shape, lazy, inner):
    """Capacities written by the build (1) or by each tile's first push pass,
    read there from the inputs, with neighbours' pushes held in per-tile
    inflow slots (2), through the single-iteration push pass (inner 1) and the
    general one: exact results, and the residual state read back (inputs +
    inflow for tiles never pushed) is a valid optimality certificate."""
    W, H = shape
    rng = np.random.default_rng(W + 7 * H + lazy)
    pos, neg, E = random_instance(rng, W, H, tmax=120, emax=40)
    o = g.default_options()
    o.lazy_caps = lazy
    o.inner_iters = inner
    gr = g.Graph(W, H, o)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    for rep in range(2):  # a rebuild of the same graph starts from clean slots
        gr.build()
        label, st = gr.optimize()
        rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
        assert st.flow == rflow and np.array_equal(label, rlabel), rep
        caps, exc = gr.read_state()
        assert ora.certify(W, H, pos, neg, caps, exc, label, st.flow, edges=E) == 0, rep
    gr.close()


def test_oversize_instance_rejected(g):
    """Instances hold at most 2^30 pixels (int32 node counts, 32-bit tile
    arithmetic); the shape is checked before anything is allocated."""
    with pytest.raises(g.GcfgError) as e:
        g.Graph(32769, 32768)
    assert e.value.rc == g.GCFG_ERR_INVALID
