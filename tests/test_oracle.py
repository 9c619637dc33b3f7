"""Pins the CPU oracle's max-flow / label semantics (CPU only).

Checks chosen so that plausible mistakes fail: the reference's own known
answer (Debug.cu:24-58), exhaustive enumeration of every cut on tiny grids
(min value AND the union of all minimum source sets, which is the unique
"cannot reach the sink" set), an independent library max-flow (networkx),
the reference border rule, and a transposition / direction check.
"""
import itertools

import numpy as np
import pytest

from conftest import debug_kat, glibc_rand_instance, random_instance


def brute_force(W, H, pos, neg, E):
    """Enumerate all labellings of the interior; return (min cut, union of optimal source sets)."""
    up, down, left, right = E
    nodes = [(x, y) for y in range(1, H - 1) for x in range(1, W - 1)]
    idx = {p: i for i, p in enumerate(nodes)}
    n = len(nodes)
    labs = np.array(list(itertools.product([0, 1], repeat=n)), dtype=np.int64)  # 2^n x n
    cost = np.zeros(len(labs), np.int64)
    for i, (x, y) in enumerate(nodes):
        s = labs[:, i]
        cost += np.where(s == 1, neg[y, x], pos[y, x])
        for (dx, dy, cap) in ((-1, 0, left), (1, 0, right), (0, -1, up), (0, 1, down)):
            q = (x + dx, y + dy)
            if q in idx:
                cost += (s == 1) * (labs[:, idx[q]] == 0) * int(cap[y, x])
    best = cost.min()
    union = labs[cost == best].max(axis=0)
    lab = np.zeros((H, W), np.int32)
    for i, (x, y) in enumerate(nodes):
        lab[y, x] = union[i]
    return int(best), lab


def test_debug_cu_known_answer(ora):
    W, H, pos, neg = debug_kat()
    flow, label = ora.maxflow(W, H, pos, neg, penalty=50)
    assert flow == 7
    assert label.sum() == 0


def test_reference_printout_emulation(ora):
    """The reference prints 4 for the Debug.cu instance (limited EnergyLevel bug, W % 32 != 0)."""
    W, H, pos, neg = debug_kat()
    flow, label, iters, printed = ora.refpr(W, H, pos, neg, penalty=50)
    assert flow == 7 and label.sum() == 0
    assert printed == 4


@pytest.mark.parametrize("seed", range(12))
def test_brute_force_tiny(ora, seed):
    rng = np.random.default_rng(seed)
    W, H = [(4, 4), (5, 4), (4, 6), (6, 5), (3, 7), (5, 5)][seed % 6]
    pos, neg, E = random_instance(rng, W, H, tmax=12, emax=8)
    # make ties likely so the union-of-optima semantics is exercised
    if seed % 2:
        pos[::2] = neg[::2]
    best, blab = brute_force(W, H, pos, neg, E)
    flow, label = ora.maxflow(W, H, pos, neg, edges=E)
    assert flow == best
    assert np.array_equal(label, blab)


def test_against_networkx(ora):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(7)
    for trial in range(20):
        W, H = int(rng.integers(3, 14)), int(rng.integers(3, 14))
        pos, neg, E = random_instance(rng, W, H, tmax=60, emax=30)
        up, down, left, right = E
        G = nx.DiGraph()
        G.add_node("s")
        G.add_node("t")
        for y in range(1, H - 1):
            for x in range(1, W - 1):
                G.add_edge("s", (x, y), capacity=int(pos[y, x]))
                G.add_edge((x, y), "t", capacity=int(neg[y, x]))
                for dx, dy, cap in ((-1, 0, left), (1, 0, right), (0, -1, up), (0, 1, down)):
                    if 1 <= x + dx <= W - 2 and 1 <= y + dy <= H - 2:
                        G.add_edge((x, y), (x + dx, y + dy), capacity=int(cap[y, x]))
        fv, (S, T) = nx.minimum_cut(G, "s", "t")
        flow, label = ora.maxflow(W, H, pos, neg, edges=E)
        assert flow == fv
        assert ora.cut_value(W, H, pos, neg, label, edges=E) == fv


def test_border_is_not_part_of_the_graph(ora):
    """Huge t-links on border pixels change nothing (GraphCutKernels.cu:122-146)."""
    rng = np.random.default_rng(1)
    W, H = 9, 7
    pos, neg, E = random_instance(rng, W, H)
    f0, l0 = ora.maxflow(W, H, pos, neg, edges=E)
    pos2, neg2 = pos.copy(), neg.copy()
    pos2[0, :] = 10 ** 6
    neg2[:, -1] = 10 ** 6
    E2 = tuple(e.copy() for e in E)
    E2[0][1, :] = 10 ** 6  # up from row 1 points into the border
    E2[1][H - 2, :] = 10 ** 6  # down from row H-2 points into the border
    f1, l1 = ora.maxflow(W, H, pos2, neg2, edges=E2)
    assert f0 == f1 and np.array_equal(l0, l1)
    assert l1[0].sum() == 0 and l1[:, -1].sum() == 0


def test_directions_are_not_transposed(ora):
    """A single strong arc in one direction: only the correct orientation carries flow."""
    W, H = 5, 3  # interior (1,1) (2,1) (3,1)
    pos = np.zeros((H, W), np.int32)
    neg = np.zeros((H, W), np.int32)
    pos[1, 1] = 10
    neg[1, 2] = 10
    z = np.zeros((H, W), np.int32)
    right = z.copy()
    right[1, 1] = 7  # (1,1) -> (2,1)
    f, lab = ora.maxflow(W, H, pos, neg, edges=(z, z, z, right))
    assert f == 7
    left = z.copy()
    left[1, 1] = 7  # (1,1) -> (0,1) is the border: no flow
    f2, _ = ora.maxflow(W, H, pos, neg, edges=(z, z, left, z))
    assert f2 == 0
    left2 = z.copy()
    left2[1, 2] = 7  # (2,1) -> (1,1): wrong direction for s->(1,1)->(2,1)->t
    f3, _ = ora.maxflow(W, H, pos, neg, edges=(z, z, left2, z))
    assert f3 == 0


def test_degenerate_and_invalid(ora):
    for W, H in ((1, 1), (2, 5), (5, 2)):
        f, lab = ora.maxflow(W, H, np.ones((H, W)), np.zeros((H, W)), penalty=1)
        assert f == 0 and lab.sum() == 0
    E = tuple(np.ones((5, 5), np.int32) for _ in range(4))
    E[1][2, 2] = -1  # negative n-links are invalid (negative t-links are not)
    with pytest.raises(RuntimeError):
        ora.maxflow(5, 5, np.zeros((5, 5)), np.zeros((5, 5)), edges=E)


def test_reference_restatement_matches_oracle(ora):
    rng = np.random.default_rng(11)
    for trial in range(6):
        W, H = int(rng.integers(3, 70)), int(rng.integers(3, 40))
        pos, neg, E = random_instance(rng, W, H, tmax=200, emax=60)
        f, l = ora.maxflow(W, H, pos, neg, edges=E)
        fr, lr, it, _ = ora.refpr(W, H, pos, neg, edges=E)
        assert fr == f and np.array_equal(lr, l)


def test_debug_cu_random_instance(ora):
    """Debug.cu:60-74 (glibc rand, seed 1, 400x400, penalty 50): oracle and the
    reference-schedule restatement agree, and the cut certifies the flow."""
    pos, neg = glibc_rand_instance(1)
    f, l = ora.maxflow(400, 400, pos, neg, penalty=50)
    assert ora.cut_value(400, 400, pos, neg, l, penalty=50) == f
    fr, lr, it, _ = ora.refpr(400, 400, pos, neg, penalty=50, threads=4)
    assert fr == f and np.array_equal(lr, l)


# ---------------------------------------------------------------------------
# 8-neighbourhood (NEIGHBORHOOD == 8, GraphCut.h:22,65-70; diagonal arcs in
# GraphCutKernels.cu:133-151,190-213).
D8 = ((-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1))


def random_instance8(rng, W, H, tmax=12, emax=8):
    pos = rng.integers(0, tmax, (H, W)).astype(np.int32)
    neg = rng.integers(0, tmax, (H, W)).astype(np.int32)
    dirs = tuple(rng.integers(0, emax, (H, W)).astype(np.int32) for _ in range(8))
    return pos, neg, dirs


def brute_force8(W, H, pos, neg, dirs):
    nodes = [(x, y) for y in range(1, H - 1) for x in range(1, W - 1)]
    idx = {p: i for i, p in enumerate(nodes)}
    labs = np.array(list(itertools.product([0, 1], repeat=len(nodes))), dtype=np.int64)
    cost = np.zeros(len(labs), np.int64)
    for i, (x, y) in enumerate(nodes):
        s = labs[:, i]
        cost += np.where(s == 1, neg[y, x], pos[y, x])
        for (dx, dy), cap in zip(D8, dirs):
            q = (x + dx, y + dy)
            if q in idx:
                cost += (s == 1) * (labs[:, idx[q]] == 0) * int(cap[y, x])
    best = cost.min()
    union = labs[cost == best].max(axis=0)
    lab = np.zeros((H, W), np.int32)
    for i, (x, y) in enumerate(nodes):
        lab[y, x] = union[i]
    return int(best), lab


@pytest.mark.parametrize("seed", range(10))
def test_brute_force_tiny8(ora, seed):
    rng = np.random.default_rng(500 + seed)
    W, H = [(4, 4), (5, 4), (4, 6), (6, 5), (3, 7)][seed % 5]
    pos, neg, dirs = random_instance8(rng, W, H)
    if seed % 2:
        pos[::2] = neg[::2]
    best, blab = brute_force8(W, H, pos, neg, dirs)
    flow, label = ora.maxflow8(W, H, pos, neg, dirs)
    assert flow == best and np.array_equal(label, blab)
    assert ora.cut_value8(W, H, pos, neg, label, dirs) == flow


def test_diagonal_only_path8(ora):
    """Source and sink joined only by a diagonal chain: every side n-link is 0,
    so the flow equals the diagonal bottleneck and a 4-neighbourhood solve is 0."""
    W, H = 7, 7
    z = np.zeros((H, W), np.int32)
    dr = z.copy()
    for k in range(1, 5):
        dr[k, k] = 5 + k  # (k,k) -> (k+1,k+1)
    dr[3, 3] = 4
    pos, neg = z.copy(), z.copy()
    pos[1, 1], neg[5, 5] = 100, 100
    dirs = (z, z, z, z, z, z, z, dr)
    flow, label = ora.maxflow8(W, H, pos, neg, dirs)
    assert flow == 4
    assert ora.maxflow(W, H, pos, neg, edges=(z, z, z, z))[0] == 0
    # the cut is at (3,3) -> (4,4): (1,1),(2,2),(3,3) are on the source side
    assert label[1, 1] == label[2, 2] == label[3, 3] == 1 and label[4, 4] == 0
    # border rule: a diagonal arc into the border carries nothing
    dirs2 = (z, z, z, z, z, z, z, np.full((H, W), 9, np.int32))
    p2, n2 = z.copy(), z.copy()
    p2[5, 5], n2[1, 1] = 50, 50  # (5,5) -> (6,6) is a border pixel
    assert ora.maxflow8(W, H, p2, n2, dirs2)[0] == 0


def test_penalty8_equals_uniform_planes(ora):
    rng = np.random.default_rng(9)
    W, H = 23, 17
    pos, neg, _ = random_instance8(rng, W, H, tmax=60)
    f1, l1 = ora.maxflow8(W, H, pos, neg, penalty=7)
    f2, l2 = ora.maxflow8(W, H, pos, neg, tuple(np.full((H, W), 7, np.int32) for _ in range(8)))
    assert f1 == f2 and np.array_equal(l1, l2)


def test_against_networkx8(ora):
    nx = pytest.importorskip("networkx")
    rng = np.random.default_rng(17)
    for trial in range(12):
        W, H = int(rng.integers(3, 12)), int(rng.integers(3, 12))
        pos, neg, dirs = random_instance8(rng, W, H, tmax=60, emax=30)
        G = nx.DiGraph()
        G.add_node("s")
        G.add_node("t")
        for y in range(1, H - 1):
            for x in range(1, W - 1):
                G.add_edge("s", (x, y), capacity=int(pos[y, x]))
                G.add_edge((x, y), "t", capacity=int(neg[y, x]))
                for (dx, dy), cap in zip(D8, dirs):
                    xx, yy = x + dx, y + dy
                    if 0 < xx < W - 1 and 0 < yy < H - 1:
                        G.add_edge((x, y), (xx, yy), capacity=int(cap[y, x]))
        want = nx.maximum_flow_value(G, "s", "t") if W > 2 and H > 2 else 0
        flow, label = ora.maxflow8(W, H, pos, neg, dirs)
        assert flow == want


# ---------------------------------------------------------------------------
# The residual-state certificate used for full-size GPU results.

def test_certificate_accepts_oracle_states_and_rejects_corruptions(ora):
    rng = np.random.default_rng(77)
    for trial in range(8):
        W, H = int(rng.integers(3, 40)), int(rng.integers(3, 30))
        pos, neg, E = random_instance(rng, W, H, tmax=80, emax=30)
        if trial % 2:
            pos = pos - rng.integers(0, 30, pos.shape).astype(np.int32)  # negative t-links too
        flow, lab, caps, exc = ora.maxflow_state(W, H, pos, neg, edges=E)
        assert ora.certify(W, H, pos, neg, caps, exc, lab, flow, edges=E) == 0
        if W < 4 or H < 4:
            continue
        # a flipped label
        bad = lab.copy()
        bad[H // 2, W // 2] ^= 1
        assert ora.certify(W, H, pos, neg, caps, exc, bad, flow, edges=E) == 3
        # a wrong flow value
        assert ora.certify(W, H, pos, neg, caps, exc, lab, flow + 1, edges=E) == 4
        # a residual that breaks the pair sum
        c2 = caps.copy()
        c2[H // 2, W // 2, 1] += 1
        assert ora.certify(W, H, pos, neg, c2, exc, lab, flow, edges=E) == 1
        # excess that breaks conservation
        e2 = exc.copy()
        e2[H // 2, W // 2] += 1
        assert ora.certify(W, H, pos, neg, caps, e2, lab, flow, edges=E) == 2


def test_certificate_rejects_a_non_maximum_preflow(ora):
    """The untouched pre-saturated graph is a feasible preflow; unless it is
    already maximum, its value is below the cut of the labels it implies."""
    rng = np.random.default_rng(5)
    W, H = 30, 20
    pos, neg, E = random_instance(rng, W, H, tmax=80, emax=30)
    up, down, left, right = E
    inter = np.zeros((H, W), bool)
    inter[1:-1, 1:-1] = True
    caps = np.zeros((H, W, 4), np.int32)
    for k, (a, (dx, dy)) in enumerate(zip((left, right, up, down), ((-1, 0), (1, 0), (0, -1), (0, 1)))):
        ys, xs = np.mgrid[0:H, 0:W]
        ok = inter & (xs + dx > 0) & (xs + dx < W - 1) & (ys + dy > 0) & (ys + dy < H - 1)
        caps[..., k] = np.where(ok, a, 0)
    exc = np.where(inter, pos - neg, 0).astype(np.int32)
    # labels: cannot reach a deficit node in this residual graph
    reach = inter & (exc < 0)
    while True:
        new = reach.copy()
        for k, (dx, dy) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            shifted = np.zeros_like(reach)
            ys, xs = np.nonzero(reach)
            us, ux = ys - dy, xs - dx
            m = (us > 0) & (us < H - 1) & (ux > 0) & (ux < W - 1)
            shifted[us[m], ux[m]] = True
            new |= shifted & (caps[..., k] > 0)
        if (new == reach).all():
            break
        reach = new
    lab = (inter & ~reach).astype(np.int32)
    value = int(pos[inter].sum() - np.maximum(exc[inter], 0).sum())
    best = ora.maxflow(W, H, pos, neg, edges=E)[0]
    assert value < best
    assert ora.certify(W, H, pos, neg, caps, exc, lab, value, edges=E) == 4
