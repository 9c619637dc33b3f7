"""Batched solve: many instances of uneven size in one engine; each result
equals the oracle and the plain drop-in on the same instance."""
import numpy as np
import pytest

from conftest import random_instance

pytestmark = pytest.mark.gpu


def _ellipse_links(g, W, H, seed):
    img = g.gen_ellipse_image(W, H, seed)
    dt = g.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(g.gen_seed_map(W, H))
    dt.compute()
    links = dt.links()
    dt.close()
    return links


@pytest.mark.parametrize("cfg", [{}, dict(gr_cap=16, gr_block_w=1, gr_block_h=4), dict(gr_cap=8, kernels=0),
                                 dict(gr_cap=0), dict(gr_cap=0, gr_block_w=2, gr_block_h=2)])
def test_mixed_batch_matches_oracle(g, ora, cfg):
    rng = np.random.default_rng(12)
    items, refs = [], []
    shapes = [(70, 40), (3, 3), (1, 9), (257, 130), (64, 16), (129, 17), (50, 300), (400, 200)]
    for i, (W, H) in enumerate(shapes):
        if i % 2:
            pos, neg, up, down, left, right = _ellipse_links(g, W, H, i + 1) if W >= 40 and H >= 40 else \
                (lambda p, n, E: (p, n) + E)(*random_instance(rng, W, H))
            E = (up, down, left, right)
            items.append(g.BatchItem(W, H, pos, neg, edges=E))
            refs.append(ora.maxflow(W, H, pos, neg, edges=E))
        else:
            pos, neg, _ = random_instance(rng, W, H, tmax=200, edges=False)
            items.append(g.BatchItem(W, H, pos, neg, penalty=37))
            refs.append(ora.maxflow(W, H, pos, neg, penalty=37))
    o = g.default_options()
    for k, v in cfg.items():
        setattr(o, k, v)
    b = g.Batch(items, o)
    for rep in range(2):  # the same batch object solves again from scratch
        flows, labels, st = b.solve(labels_to_host=True)
        for k, (rf, rl) in enumerate(refs):
            assert flows[k] == rf, (rep, k)
            assert np.array_equal(labels[k], rl), (rep, k)
    flows_d, _, _ = b.solve(labels_to_host=False)
    assert np.array_equal(flows_d, flows)
    for k, (rf, rl) in enumerate(refs):
        assert np.array_equal(b.device_labels(k), rl)
    # Packed labels on the host: the same maps, bits past W zero. Before the
    # second solve the pinned buffer is overwritten, so every instance's
    # words must be written again by that solve.
    for rep in range(2):
        if rep:
            for w in b.solve_bits(copy=False)[1]:
                w.fill(~np.uint64(0))
        flows_b, words, _ = b.solve_bits()
        assert np.array_equal(flows_b, flows), rep
        for k, (rf, rl) in enumerate(refs):
            W, H = shapes[k]
            assert words[k].shape == (H, (W + 63) // 64)
            assert np.array_equal(words[k], g.pack_bits(rl)), (rep, k)
    b.close()


def test_batch_of_ellipse_images_equals_plain(g, ora):
    Ws, Hs = g.gen_batch_shapes(24, 5, 128.0 ** 2, 512.0 ** 2)
    items, plain = [], []
    for i in range(len(Ws)):
        W, H = int(Ws[i]), int(Hs[i])
        pos, neg, up, down, left, right = _ellipse_links(g, W, H, g.mix_seed(5, i))
        items.append(g.BatchItem(W, H, pos, neg, edges=(up, down, left, right)))
        flow, label, _ = g.solve(W, H, pos, neg, edges=(up, down, left, right))
        plain.append((flow, label))
        if i < 4:
            rf, rl = ora.maxflow(W, H, pos, neg, edges=(up, down, left, right))
            assert rf == flow and np.array_equal(rl, label)
    b = g.Batch(items)
    flows, labels, _ = b.solve()
    for k, (f, l) in enumerate(plain):
        assert flows[k] == f and np.array_equal(labels[k], l)
    b.close()


def test_autotune_keeps_results(g, ora):
    rng = np.random.default_rng(8)
    items, refs = [], []
    for i, (W, H) in enumerate([(130, 70), (64, 64), (200, 90)]):
        pos, neg, E = random_instance(rng, W, H, tmax=150, emax=40)
        items.append(g.BatchItem(W, H, pos, neg, edges=E))
        refs.append(ora.maxflow(W, H, pos, neg, edges=E))
    b = g.Batch(items)
    best, ms = b.autotune(trials=1)
    assert ms > 0 and best.inner_iters >= 1 and best.sweeps_per_gr >= 1
    flows, labels, _ = b.solve()
    for k, (rf, rl) in enumerate(refs):
        assert flows[k] == rf and np.array_equal(labels[k], rl)
    b.close()
    W, H = 150, 80
    pos, neg, E = random_instance(rng, W, H, tmax=150, emax=40)
    gr = g.Graph(W, H)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    best, ms = gr.autotune(trials=1)
    gr.build()
    label, st = gr.optimize()
    rf, rl = ora.maxflow(W, H, pos, neg, edges=E)
    assert st.flow == rf and np.array_equal(label, rl)
    gr.close()


def _stress_links(g, W, H, seed):
    """Heavy-tailed t-links and Perlin contrast n-links: instances that need
    more than the four capped relabel rounds, so exact and partial exact
    rounds run."""
    pos, neg = g.gen_stress_tlinks(W, H, seed)
    dt = g.DataTerm(W, H)
    dt.set_image(g.gen_perlin_image(W, H, seed + 1))
    dt.set_seeds(np.zeros((H, W), np.uint8))
    dt.compute()
    _, _, up, down, left, right = dt.links()
    dt.close()
    return pos, neg, (up, down, left, right)


@pytest.mark.parametrize("lazy", [0, 1, 2])
def test_batch_with_partial_rounds(g, ora, lazy):
    """Hard instances (many relabel rounds: growing capped rounds, exact
    rounds that recompute only the blocks near the last pushes, a full round
    at the end) next to easy ones in one batch, with the capacities written by
    the build, lazily, or chosen per build: every flow and label map equals
    the oracle's, on two solves of the same batch."""
    items, refs = [], []
    for i, (W, H) in enumerate([(640, 480), (300, 200), (200, 700), (320, 240), (129, 77), (512, 512)]):
        if i % 3 == 0 or i == 5:
            pos, neg, E = _stress_links(g, W, H, 40 + i)
        else:
            pos, neg, up, down, left, right = _ellipse_links(g, W, H, 50 + i)
            E = (up, down, left, right)
        items.append(g.BatchItem(W, H, pos, neg, edges=E))
        refs.append(ora.maxflow(W, H, pos, neg, edges=E))
    o = g.default_options()
    o.lazy_caps = lazy
    b = g.Batch(items, o)
    for rep in range(2):
        flows, labels, st = b.solve(labels_to_host=True)
        for k, (rf, rl) in enumerate(refs):
            assert flows[k] == rf, (rep, k)
            assert np.array_equal(labels[k], rl), (rep, k, int((labels[k] != rl).sum()))
    assert st.gr_rounds > 6 * 5  # summed over instances: the hard ones ran past their capped rounds
    b.close()
