"""Parity at the full sizes of the benchmarked configurations:
  * 1024 x 1024 ellipse image through the device data term and the plain
    GC_* path: flow and labels equal the oracle's, and the residual state
    passes the oracle's optimality certificate;
  * 4096 x 4096 heavy-tailed t-links on 4-octave Perlin n-links (too large to
    re-solve on the CPU in a test): the residual state passes the certificate
    (feasible preflow, conservation, labels = nodes that cannot reach a
    deficit node, flow = sum pos - sum max(excess, 0) = cut of the labels);
  * the 256-image batch of areas 128^2..1024^2: every instance equals the
    plain path and the oracle bit for bit, with the default options, the
    benchmark's tuned options and both capacity layouts, each solved twice."""
import os

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


@pytest.fixture(scope="module")
def c3_batch(g, ora):
    """The 256-image benchmark batch (device data term) with every instance's
    oracle flow and labels (solved on the host cores in parallel)."""
    from concurrent.futures import ThreadPoolExecutor

    n = 256
    Ws, Hs = g.gen_batch_shapes(n, 2024, 128.0 ** 2, 1024.0 ** 2)
    print(f"[fullsize] batch of {n}, {int((Ws * Hs).sum())} px", flush=True)
    host = []
    for i in range(n):
        W, H = int(Ws[i]), int(Hs[i])
        host.append((W, H, _links(g, W, H, g.gen_ellipse_image(W, H, g.mix_seed(2024, i)), g.gen_seed_map(W, H))))

    def ref(h):
        W, H, (pos, neg, up, down, left, right) = h
        return ora.maxflow(W, H, pos, neg, edges=(up, down, left, right))

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        refs = list(ex.map(ref, host))
    return host, refs


def test_c3_batch_256(g, ora, c3_batch):
    host, refs = c3_batch
    items = [g.BatchItem(W, H, L[0], L[1], edges=tuple(L[2:6])) for W, H, L in host]
    b = g.Batch(items)
    flows, labels, _ = b.solve(labels_to_host=True)
    b.close()
    for i, (W, H, (pos, neg, up, down, left, right)) in enumerate(host):
        f, lab, _ = g.solve(W, H, pos, neg, edges=(up, down, left, right))
        assert flows[i] == f and np.array_equal(labels[i], lab), i
        rf, rl = refs[i]
        assert f == rf and np.array_equal(lab, rl), i


# The options the benchmark's tuner picks for this batch (one push iteration
# per pass, one sweep pair per relabel, cap 4, 56 relabel wavefronts per CU),
# with the capacity layout left to the engine, forced to the build, and forced
# lazy. Each option set solves the batch twice on one handle: the second
# solve starts with the first solve's step count as its first chunk, which
# may switch the capacity layout and changes the poll schedule.
C3_OPTION_SETS = [dict(), dict(sweeps_per_gr=1, inner_iters=1, gr_cap=4, gr_waves_per_cu=56),
                  dict(sweeps_per_gr=1, inner_iters=1, gr_cap=4, gr_waves_per_cu=56, lazy_caps=1),
                  dict(sweeps_per_gr=1, inner_iters=1, gr_cap=4, gr_waves_per_cu=56, lazy_caps=2),
                  dict(gr_cap=8, lazy_caps=1), dict(gr_cap=8, lazy_caps=2)]


@pytest.mark.parametrize("opts", C3_OPTION_SETS)
def test_c3_batch_every_instance_vs_oracle(g, c3_batch, opts):
    host, refs = c3_batch
    o = g.default_options()
    for k, v in opts.items():
        setattr(o, k, v)
    items = [g.BatchItem(W, H, L[0], L[1], edges=tuple(L[2:6])) for W, H, L in host]
    b = g.Batch(items, o)
    for rep in range(3):
        print(f"[fullsize] batch {opts} solve {rep}", flush=True)
        if rep < 2:
            flows, labels, _ = b.solve(labels_to_host=True)
        else:  # labels packed one bit per pixel on the host
            flows, words, _ = b.solve_bits()
            labels = [g.unpack_bits(w, W, H) for w, (W, H, _) in zip(words, host)]
        for i, (rf, rl) in enumerate(refs):
            assert flows[i] == rf, (opts, rep, i, int(flows[i]), rf)
            assert np.array_equal(labels[i], rl), (opts, rep, i, int((labels[i] != rl).sum()))
    b.close()
