"""Parity at the full sizes of the benchmarked configurations:
  * 1024 x 1024 ellipse image through the device data term and the plain
    GC_* path: flow and labels equal the oracle's, and the residual state
    passes the oracle's optimality certificate;
  * 4096 x 4096 heavy-tailed t-links on 4-octave Perlin n-links (too large to
    re-solve on the CPU in a test): the residual state passes the certificate
    (feasible preflow, conservation, labels = nodes that cannot reach a
    deficit node, flow = sum pos - sum max(excess, 0) = cut of the labels);
  * the 256-image batch of areas 128^2..1024^2: every instance equals the
    plain path bit for bit, and a sample equals the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _links(g, W, H, img, seeds):
    dt = g.DataTerm(W, H)
    dt.set_image(img)
    dt.set_seeds(seeds)
    dt.compute()
    out = dt.links()
    dt.close()
    return out


def _solve_state(g, W, H, pos, neg, E, options=None):
    gr = g.Graph(W, H, options)
    gr.set_dataterms(pos, neg)
    gr.set_edges(*E)
    gr.build()
    label, st = gr.optimize()
    caps, exc = gr.read_state()
    gr.close()
    return label, st, caps, exc


def test_c1_ellipse_1024(g, ora):
    W = H = 1024
    print("[fullsize] 1024^2 ellipse", flush=True)
    pos, neg, up, down, left, right = _links(g, W, H, g.gen_ellipse_image(W, H, 1), g.gen_seed_map(W, H))
    E = (up, down, left, right)
    label, st, caps, exc = _solve_state(g, W, H, pos, neg, E)
    rflow, rlabel = ora.maxflow(W, H, pos, neg, edges=E)
    assert st.flow == rflow and np.array_equal(label, rlabel)
    assert ora.certify(W, H, pos, neg, caps, exc, label, st.flow, edges=E) == 0


@pytest.mark.parametrize("cap", [0, 8])
def test_c2_stress_4096_certificate(g, ora, cap):
    W = H = 4096
    print(f"[fullsize] 4096^2 heavy-tailed, gr_cap {cap}", flush=True)
    _, _, up, down, left, right = _links(g, W, H, g.gen_perlin_image(W, H, 2024), np.zeros((H, W), np.uint8))
    pos, neg = g.gen_stress_tlinks(W, H, 2025)
    o = g.default_options()
    o.gr_cap = cap
    E = (up, down, left, right)
    label, st, caps, exc = _solve_state(g, W, H, pos, neg, E, o)
    assert st.gr_rounds > 1 and 0 < label.sum() < (W - 2) * (H - 2)
    assert ora.certify(W, H, pos, neg, caps, exc, label, st.flow, edges=E) == 0


def test_c3_batch_256(g, ora):
    n = 256
    Ws, Hs = g.gen_batch_shapes(n, 2024, 128.0 ** 2, 1024.0 ** 2)
    print(f"[fullsize] batch of {n}, {int((Ws * Hs).sum())} px", flush=True)
    items, host = [], []
    for i in range(n):
        W, H = int(Ws[i]), int(Hs[i])
        links = _links(g, W, H, g.gen_ellipse_image(W, H, g.mix_seed(2024, i)), g.gen_seed_map(W, H))
        host.append((W, H, links))
        items.append(g.BatchItem(W, H, links[0], links[1], edges=tuple(links[2:6])))
    b = g.Batch(items)
    flows, labels, _ = b.solve(labels_to_host=True)
    b.close()
    for i, (W, H, (pos, neg, up, down, left, right)) in enumerate(host):
        f, lab, _ = g.solve(W, H, pos, neg, edges=(up, down, left, right))
        assert flows[i] == f and np.array_equal(labels[i], lab), i
        if i % 32 == 0:
            rf, rl = ora.maxflow(W, H, pos, neg, edges=(up, down, left, right))
            assert f == rf and np.array_equal(lab, rl), i
