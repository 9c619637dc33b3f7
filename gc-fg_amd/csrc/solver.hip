// CDNA4 kernels of the exact push-relabel min-cut solver.
//
// Replaces GraphCutKernels.cu (InitGraph*/Relabel/WavePush/UpdateActivity/
// InitGlobalRelabel/GlobalRelabel/InitLabels/SpreadLabels/EnergyLevel,
// :112-898) with a race-free design:
//  * k_build: residual graph from caller inputs, border rule of
//    GraphCutKernels.cu:122-146 / :179-203, t-links pre-saturated
//    (excess = pos - neg, :197), interior sum of pos in int64.
//  * k_pr: one 64x16 tile per workgroup, state in registers + LDS, up to
//    inner_iters synchronous iterations of {push along admissible arcs,
//    receive, relabel}; tiles run in a 2-colour checkerboard so the halo
//    heights a tile reads are stable for the whole pass (every pass is a
//    legal sequential push-relabel schedule, so the labelling stays valid).
//    Flow into halo nodes is applied with int32 atomics at the end.
//  * k_gr: exact BFS distance to the sink-connected nodes (excess < 0) by
//    tile-local relaxation to a fixed point in LDS, tiles re-queued through
//    dirty flags until no boundary changes (replaces the approximate,
//    host-polled GlobalRelabel of GraphCutKernels.cu:757-875), plus FIN:
//    labels (interior and unreachable, the SPREAD_ZEROS set of :228-351)
//    and the int64 stuck-excess sum for flow = sum pos - sum max(e, 0).
//  * k_sched: per-instance state machine and ballot-compacted tile queues.
#include <hip/hip_ext.h>

#include <atomic>
#include <type_traits>

#include "engine.hpp"

namespace gcfg {

constexpr int GR_MAX_ITERS = 4 * (TW + TH);

// Caller input planes and label buffers are plain device memory, reached
// through the descriptor's generic pointers: cast them to the global address
// space so that the accesses are global (not flat) loads and stores, which do
// not also count against the LDS/scalar counter and force wider waits.
__device__ __forceinline__ int ldg_in(const int32_t *p, size_t i) {
    return ((const __attribute__((address_space(1))) int32_t *)p)[i];
}
__device__ __forceinline__ void stg_out(int32_t *p, size_t i, int v) {
    ((__attribute__((address_space(1))) int32_t *)p)[i] = v;
}
__device__ __forceinline__ void stg_out64(uint64_t *p, size_t i, uint64_t v) {
    ((__attribute__((address_space(1))) uint64_t *)p)[i] = v;
}

// FIN of the workgroup kernels (k_gr, k_gr8): labels of the RPT rows that
// one wavefront owns (lane = column) from the final distances, as int32 per
// pixel and/or one packed word per row (InstDesc::label_bits), and the
// wavefront's share of the tile's stuck excess.
__device__ __forceinline__ int64_t fin_rows_wg(const KParams &p, const InstDesc &D, const TileInfo &ti,
                                               size_t base, int wave, int lane) {
    int64_t sx = 0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int y = wave * RPT + k;
        const int X = ti.tx * TW + lane, Y = ti.ty * TH + y;
        const size_t g = base + y * TW + lane;
        const bool in = X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1;
        const int lab = (in && p.dist[g] >= DINF) ? 1 : 0;
        if (D.label_out && X < D.W && Y < D.H) stg_out(D.label_out, (size_t)Y * D.W + X, lab);
        if (in) sx += max(p.exc[g], 0);
        if (D.label_bits) {
            const uint64_t b = __ballot(lab);
            if (lane == 0 && Y < D.H) stg_out64(D.label_bits, (size_t)Y * D.tiles_x + ti.tx, b);
        }
    }
    return sx;
}

// Residual capacities {left, right, up, down} of pixel (X, Y) as the build
// defines them from the inputs: zero outside the interior and on arcs into the
// border (GraphCutKernels.cu:122-146, :179-203).
// (Loaded from a clamped in-bounds address and masked, not branched around;
// an in-bounds index is below 2^30, shapes being at most 2^30 pixels, and a
// 32-bit product keeps the selection a select.)
__device__ __forceinline__ int4 input_caps(const InstDesc &D, int X, int Y) {
    const bool in = (X > 0) & (Y > 0) & (X < D.W - 1) & (Y < D.H - 1);
    const size_t i = in ? (size_t)((uint32_t)Y * (uint32_t)D.W + (uint32_t)X) : 0;
    int4 c;
    if (D.up)
        c = make_int4(ldg_in(D.left, i), ldg_in(D.right, i), ldg_in(D.up, i), ldg_in(D.down, i));
    else
        c = make_int4(D.penalty, D.penalty, D.penalty, D.penalty);
    c.x = (in && X != 1) ? c.x : 0;
    c.y = (in && X != D.W - 2) ? c.y : 0;
    c.z = (in && Y != 1) ? c.z : 0;
    c.w = (in && Y != D.H - 2) ? c.w : 0;
    return c;
}

// Pending reverse-arc capacity of tile-local node (x, y) (see KParams::inflow).
__device__ __forceinline__ int4 inflow_of(const KParams &p, int t, int x, int y) {
    const int32_t *f = p.inflow + (size_t)t * INFLOW;
    return make_int4(x == 0 ? f[y] : 0, x == TW - 1 ? f[TH + y] : 0, y == 0 ? f[2 * TH + x] : 0,
                     y == TH - 1 ? f[2 * TH + TW + x] : 0);
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int lo = __shfl_xor((int)(uint32_t)(uint64_t)v, o, 64);
        const int hi = __shfl_xor((int)(uint32_t)((uint64_t)v >> 32), o, 64);
        v += (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
    }
    return v;
}

// Block (NT threads) sum; result valid in every thread. Uses red[NT/64].
template <int NTHR = NT>
__device__ __forceinline__ int64_t block_sum64(int64_t v, int64_t *red) {
    v = wave_sum64(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    int64_t s = 0;
#pragma unroll
    for (int w = 0; w < NTHR / 64; ++w) s += red[w];
    return s;
}

__device__ __forceinline__ void atomic_add64(int64_t *p, int64_t v) {
    atomicAdd(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v);
}

// Tile flags are plain stores: the scheduler scans them between launches, so
// concurrent writers of the same value need no atomics.
__device__ __forceinline__ void set_active(const KParams &p, int t, int v) { p.act_flag[t] = v; }

// Inputs a node accepts: t-links of any sign, as in the reference (it takes
// excess = pos - neg unchecked, GraphCutKernels.cu:196-197; its own colour
// data term makes neg negative where p > 1, DataTermKernels.cu:81-82), with
// |pos - neg| <= 2^29, and n-links in [0, 2^28) (4-neighbourhood) or
// [0, 2^27) (8-neighbourhood). Excess then stays inside int32 for good: with
// e0 = pos - neg plus the accumulated edits of a session (at most 2^29, kept
// by the server), a node's excess is e0 plus the net flow of its pairs, each
// at most the pair's capacity into the node, so |e| < 2^29 + 2^29 + 2^30;
// and a pair's residual capacity is at most the sum of its two capacities.
constexpr int TLINK_DIFF_MAX = 1 << 29;
__device__ __forceinline__ bool tlinks_bad(int pp, int qq) {
    const long long d = (long long)pp - qq;
    return d < -(long long)TLINK_DIFF_MAX || d > TLINK_DIFF_MAX;
}
__device__ __forceinline__ bool nlink_bad(int c, int nbhd) { return c < 0 || c >= (nbhd == 8 ? 1 << 27 : 1 << 28); }
__device__ __forceinline__ bool nlinks_bad(const int4 &c, int nbhd) {
    return nlink_bad(c.x, nbhd) | nlink_bad(c.y, nbhd) | nlink_bad(c.z, nbhd) | nlink_bad(c.w, nbhd);
}
__device__ __forceinline__ void mark_dirty(const KParams &p, int t) { p.dirty[t] = 1; }

// ---------------------------------------------------------------------------
// Arm: reset the per-instance state and mark it live.
__device__ void clear_tile(const KParams &p, const int32_t *slots, const int32_t *prefix, int n, int b);

// Arms the listed instances (thread i: slots[i]), clears the iteration
// counters and, before a wavefront build, the first `clear` tile records of
// the listed instances (build-list positions, prefix = tile offsets).
__global__ void k_arm(KParams p, const int32_t *slots, int n, int warm, int sweeps, int start_phase, int clear) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        p.ctl->pr_iters = 0;
        p.ctl->gr_iters = 0;
    }
    for (int b = i; b < clear; b += gridDim.x * blockDim.x) clear_tile(p, slots, slots + n, n, b);
    if (i >= n) return;
    const int s = slots[i];
    InstState &st = p.state[s];
    const int64_t sp = warm ? st.sum_pos : 0;
    const int was_live = st.phase != PH_IDLE && st.phase != PH_DONE;
    InstState z{};
    z.sum_pos = sp;
    z.phase = start_phase;
    z.gr_rounds = start_phase == PH_GR ? 1 : 0;
    z.sweeps = sweeps;
    // The wavefront build runs the first (possibly capped) relabel sweep.
    z.cap = (start_phase == PH_GR && p.gr_cap > 0) ? p.gr_cap : DINF;
    st = z;
    if (!was_live) atomicAdd(&p.ctl->n_live, 1);  // re-arming a live instance keeps one count
}

// ---------------------------------------------------------------------------
// Build (GC_SetGraph): one tile per loop trip over the listed instances.
__global__ __launch_bounds__(NT) void k_build(KParams p, const int32_t *slots, const int32_t *prefix, int n,
                                              int total) {
    __shared__ int64_t red[NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int b = blockIdx.x; b < total; b += gridDim.x) {
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
        }
        const int s = slots[lo];
        const InstDesc D = p.desc[s];
        const int lt = b - prefix[lo];
        const int t = D.tile0 + lt;
        const int tx = lt % D.tiles_x, ty = lt / D.tiles_x;
        if (threadIdx.x == 0) {
            TileInfo ti;
            ti.inst = s;
            ti.tx = tx;
            ti.ty = ty;
            ti.nbr[0] = tx > 0 ? t - 1 : -1;
            ti.nbr[1] = tx < D.tiles_x - 1 ? t + 1 : -1;
            ti.nbr[2] = ty > 0 ? t - D.tiles_x : -1;
            ti.nbr[3] = ty < D.tiles_y - 1 ? t + D.tiles_x : -1;
            ti.pad = 0;
            p.tiles[t] = ti;
            if (p.nbhd == 8) {
                const bool l = tx > 0, r = tx < D.tiles_x - 1, u = ty > 0, d = ty < D.tiles_y - 1;
                p.nbr8[t] = make_int4(l && u ? t - D.tiles_x - 1 : -1, r && u ? t - D.tiles_x + 1 : -1,
                                      l && d ? t + D.tiles_x - 1 : -1, r && d ? t + D.tiles_x + 1 : -1);
            }
            p.act_flag[t] = 0;
            p.dirty[t] = 0;
            p.live[t] = 2;
            p.stale[t] = 0;
        }
        int64_t sp = 0;
        int bad = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k, x = lane;
            const int X = tx * TW + x, Y = ty * TH + y;
            const size_t g = (size_t)t * TN + y * TW + x;
            int4 c = make_int4(0, 0, 0, 0), c8 = make_int4(0, 0, 0, 0);
            int e = 0;
            if (X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1) {
                const size_t i = (size_t)Y * D.W + X;
                const int pp = D.pos[i], qq = D.neg[i];
                if (p.nbhd == 8) {
                    // Diagonal arcs (GraphCutKernels.cu:133-151,190-213): zero
                    // towards the border like the side arcs.
                    if (D.up)
                        c8 = make_int4(D.upleft[i], D.upright[i], D.downleft[i], D.downright[i]);
                    else
                        c8 = make_int4(D.penalty, D.penalty, D.penalty, D.penalty);
                    if (X == 1) c8.x = c8.z = 0;
                    if (X == D.W - 2) c8.y = c8.w = 0;
                    if (Y == 1) c8.x = c8.y = 0;
                    if (Y == D.H - 2) c8.z = c8.w = 0;
                    bad |= nlinks_bad(c8, 8);
                }
                if (D.up) {
                    c.x = D.left[i];
                    c.y = D.right[i];
                    c.z = D.up[i];
                    c.w = D.down[i];
                } else {
                    c = make_int4(D.penalty, D.penalty, D.penalty, D.penalty);
                }
                if (X == 1) c.x = 0;
                if (X == D.W - 2) c.y = 0;
                if (Y == 1) c.z = 0;
                if (Y == D.H - 2) c.w = 0;
                bad |= tlinks_bad(pp, qq) | nlinks_bad(c, p.nbhd);
                e = pp - qq;
                sp += pp;
            }
            p.caps[g] = c;
            p.exc[g] = e;
            p.dist[g] = DINF;
            p.mask[g] = 0;
            if (p.nbhd == 8) {
                p.caps8[g] = c8;
                p.mask8[g] = 0;
            }
        }
        // Capacities are written here: the tile is materialised, nothing pending.
        for (int k = threadIdx.x; k < INFLOW; k += NT) p.inflow[(size_t)t * INFLOW + k] = 0;
        const int64_t tot = block_sum64(sp, red);
        const int anybad = __syncthreads_or(bad);
        if (threadIdx.x == 0) {
            p.tsum[2 * (size_t)t] = tot;
            p.mat[t] = 1;
            if (anybad) atomicOr(&p.state[s].err, 1);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Scheduler: one workgroup per instance slot. Decisions are block-uniform
// (taken from block-wide scans of the tile flags); thread 0 writes the state.
enum EmitKind : int { EM_ALL = 0, EM_DIRTY = 1, EM_COLOUR = 2, EM_LIVE = 3, EM_DEAD = 4 };

// Per-block liveness (KParams::live) as the last relabel sweep of the block
// left it: 0 every node unreachable, 1 all at or beyond the cap, 2 some node
// below the cap, 3 (capped and reachability rounds) every interior node
// reaches a sink -- below the cap, or at 0 in a reachability round. A block
// at 3 that no push has made stale since stays at 3: a path shorter than the
// cap changes only through a push within the cap of the block.
// Append the instance's tiles (push/relabel queue) or global-relabel blocks
// (identified by their anchor tile, whose flags stand for the block)
// selected by `kind` to a queue (ballot compaction, one atomic per 256
// entries); returns the number appended.
// The scheduler runs one workgroup of SNT threads per instance slot, a pass
// of emit_tiles covering SNT * EMIT_U tiles with one queue reservation:
// 256 threads (8 units each, 2 k tiles a pass) while every instance has at
// most 2 k tiles, 1024 (16 units each, 16 k a pass: a 4096^2 image in one)
// otherwise.
template <int SNT>
constexpr int emit_units() { return SNT >= 1024 ? 16 : 8; }
constexpr int CAP_ROUNDS = 4;  // capped relabel rounds before an instance switches to exact ones
// Push-queue item flag: the tile's capacities are still only in the inputs
// (lazy build); the push pass reads them from there and writes them out.
constexpr int PR_FRESH = 1 << 30;

template <int SNT>
__device__ __forceinline__ int emit_tiles(const KParams &p, const InstDesc &D, int kind, int colour, bool to_pr,
                                          int grop, int parity) {
    constexpr int EMIT_U = emit_units<SNT>();
    __shared__ int wcount[SNT / 64], wbase[SNT / 64], s_qbase, s_total;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t *queue = to_pr ? p.queue_pr : p.queue_gr;
    int32_t *qlen = to_pr ? &p.ctl->q[parity].qlen_pr : &p.ctl->q[parity].qlen_gr;
    const int bx = to_pr ? 1 : p.gbx, by = to_pr ? 1 : p.gby;
    const bool unit = bx == 1 && by == 1;
    const int nbx = (D.tiles_x + bx - 1) / bx;
    const int nunits = unit ? D.ntiles : nbx * ((D.tiles_y + by - 1) / by);
    // The flag a unit's predicate reads first, and the second one if any
    // (push items of a lazy instance: whether the tile's capacities are
    // materialised yet).
    const int32_t *f0 = kind == EM_DIRTY ? p.dirty : kind == EM_COLOUR ? p.act_flag : p.live;
    const int32_t *f1 = kind == EM_LIVE                ? p.stale
                        : (kind == EM_DIRTY && colour) ? p.live  // reachability rounds skip live-3 blocks
                        : (kind == EM_COLOUR && D.lazy) ? p.mat
                                                        : nullptr;
    __syncthreads();  // earlier readers of s_total are done
    if (threadIdx.x == 0) s_total = 0;
    if (nunits == 0) __syncthreads();
    // Passes of SNT * EMIT_U units: wave w takes the contiguous units
    // [base + w * 64 * EMIT_U, + 64 * EMIT_U), so the queue keeps the units in
    // order, and one queue reservation covers the pass. Every flag is loaded
    // before any is cleared, so the loads of a pass are in flight together.
    for (int base = 0; base < nunits; base += SNT * EMIT_U) {
        int lt[EMIT_U], v0[EMIT_U], v1[EMIT_U];
#pragma unroll
        for (int k = 0; k < EMIT_U; ++k) {
            const int lu = base + (wave * EMIT_U + k) * 64 + lane;
            lt[k] = -1;
            v0[k] = v1[k] = 0;
            if (lu < nunits) {
                lt[k] = unit ? lu : (lu / nbx) * by * D.tiles_x + (lu % nbx) * bx;
                const int t = D.tile0 + lt[k];
                if (kind != EM_ALL) v0[k] = f0[t];
                if (f1) v1[k] = f1[t];
            }
        }
        uint64_t bal[EMIT_U];
        int wsum = 0;
#pragma unroll
        for (int k = 0; k < EMIT_U; ++k) {
            int pred = 0;
            if (lt[k] >= 0) {
                const int t = D.tile0 + lt[k];
                if (kind == EM_ALL) {
                    pred = 1;
                } else if (kind == EM_DEAD) {
                    // FIN: blocks the last round did not visit (every node
                    // unreachable, or every interior node reaching a sink)
                    pred = v0[k] == 0 || v0[k] == 3;
                } else if (kind == EM_LIVE) {
                    // Exact round (colour 1): every block with a reachable node.
                    // Reachability round (colour 2): every block with a node
                    // that may not reach a sink. Capped round: blocks with a
                    // node below the cap that a push pass may have changed
                    // (the others keep their labels).
                    // Partial exact round (colour 3): every block with a
                    // reachable node within the stale distance of a push.
                    const int lv = v0[k];
                    pred = colour == 1   ? lv != 0
                           : colour == 2 ? (lv == 1 || lv == 2)
                           : colour == 3 ? (lv != 0 && v1[k])
                                         : (lv >= 2 && v1[k]);
                    if (v1[k]) p.stale[t] = 0;
                    if (!colour && lv == 1) {
                        // A skipped block's nodes are all frozen: none is active.
                        const int tx0 = lt[k] % D.tiles_x, ty0 = lt[k] / D.tiles_x;
                        for (int j = 0; j < by && ty0 + j < D.tiles_y; ++j)
                            for (int i = 0; i < bx && tx0 + i < D.tiles_x; ++i)
                                p.act_flag[t + j * D.tiles_x + i] = 0;
                    }
                } else if (kind == EM_DIRTY) {
                    // colour 1 (a reachability round): a block whose nodes all
                    // reach a sink already cannot change, whatever its halo does.
                    pred = v0[k] && !(colour && v1[k] == 3);
                    if (v0[k]) p.dirty[t] = 0;
                } else {
                    // 4-neighbourhood: checkerboard (no two side neighbours
                    // run together). 8-neighbourhood: (tx & 1, ty & 1) four
                    // colours, so diagonal neighbours do not run together either.
                    const int tx = lt[k] % D.tiles_x, ty = lt[k] / D.tiles_x;
                    const int c = p.nbhd == 8 ? ((tx & 1) | ((ty & 1) << 1)) : ((tx + ty) & 1);
                    pred = c == colour && v0[k];
                }
            }
            bal[k] = __ballot(pred);
            wsum += __popcll(bal[k]);
        }
        if (lane == 0) wcount[wave] = wsum;
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int w = 0; w < SNT / 64; ++w) {
                wbase[w] = tot;
                tot += wcount[w];
            }
            if (p.n_slots == 1) {
                // One instance slot: this workgroup is the step's only writer
                // of the queue, which the previous step's scheduler left
                // empty, and no queue gets items from two emissions of one
                // step; the reservation is a plain store, not an atomic
                // round trip.
                s_qbase = s_total;
                if (tot) *qlen = s_total + tot;
            } else {
                s_qbase = tot ? atomicAdd(qlen, tot) : 0;
            }
            s_total += tot;
        }
        __syncthreads();
        int pos = s_qbase + wbase[wave];
        const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
        for (int k = 0; k < EMIT_U; ++k) {
            if ((bal[k] >> lane) & 1ull) {
                const int t = D.tile0 + lt[k];
                const int fresh = (kind == EM_COLOUR && D.lazy && !v1[k]) ? PR_FRESH : 0;
                queue[pos + __popcll(bal[k] & below)] = to_pr ? (t | fresh) : ((t << 3) | grop);
            }
            pos += __popcll(bal[k]);
        }
        __syncthreads();
    }
    return s_total;
}

template <int SNT>
__device__ __forceinline__ bool any_active_tile(const KParams &p, const InstDesc &D) {
    int any = 0;
#pragma unroll 8
    for (int lt = threadIdx.x; lt < D.ntiles; lt += SNT) any |= p.act_flag[D.tile0 + lt];
    return __syncthreads_or(any) != 0;
}

// A new BFS round. After a build or an edit every tile derives its mask from
// capacities and excess (OP_GR_INIT). Later rounds of the wavefront kernels
// start from the masks the push passes keep current (OP_GR_REINIT) and skip
// tiles whose nodes all had infinite distance in the previous round: a node
// that cannot reach a node with excess < 0 never can again (a push only
// creates a residual arc towards a node that can reach the sink, and sink
// nodes only lose deficit), so such a tile's distances stay infinite, no flow
// enters or leaves it, and its stored distances are already the answer.
//
// With a cap K (wavefront kernels, gr_cap > 0) a relabel computes min(d, K)
// instead of d, starting every non-sink node at K, so BFS fronts stop after K
// levels. min(d, K) is still a valid labelling; nodes at K are frozen (not
// active) until the next relabel. Only an exact relabel (cap = DINF) can end
// the solve, so a capped round that leaves no active node is followed by an
// exact one.
template <int SNT>
__device__ __forceinline__ void start_global_relabel(const KParams &p, const InstDesc &D, InstState &st,
                                                     const InstState &S0, int parity, bool reinit, int cap,
                                                     bool all = false, bool partial = false, bool grow = false,
                                                     int wcap = 0) {
    if (threadIdx.x == 0) {
        st.gr_rounds = S0.gr_rounds + 1;
        st.phase = PH_GR;
        st.cap = cap;
        st.partial = partial;
        if (cap == CAP_FLOOD) st.wcap = wcap;
    }
    // An exact round (cap DINF) revisits every block with a reachable node. A
    // capped round with cap K recomputes only blocks that have a node below
    // K and lie within K of a tile a push pass changed since the last round.
    // Exact distances never decrease, so a block with all nodes at >= K
    // stays there (its labels >= K are frozen and no push can enter such a
    // node), and for a block farther than K from every change each shortest
    // path of length < K is unchanged, so min(d, K) is unchanged too.
    const bool whole = cap >= DINF || cap == CAP_FLOOD;  // exact distances or reachability: every live block
    const int n = !reinit ? emit_tiles<SNT>(p, D, EM_ALL, 0, false, OP_GR_INIT, parity)
                  : all   ? emit_tiles<SNT>(p, D, EM_ALL, 0, false, OP_GR_REINIT, parity)
                          : emit_tiles<SNT>(p, D, EM_LIVE, cap == CAP_FLOOD ? 2 : partial ? 3 : ((whole || grow) ? 1 : 0),
                                       false, OP_GR_REINIT, parity);
    if (threadIdx.x == 0) {
        st.gr_tiles = S0.gr_tiles + n;
        st.gr_init_tiles = S0.gr_init_tiles + n;
        if (whole) st.gr_exact_tiles = S0.gr_exact_tiles + n;
    }
}

// Packed label words of a finished instance to its pinned host buffer
// (writes over PCIe from the workgroup: 16 bytes a lane, four loads in
// flight before their stores).
template <int SNT>
__device__ __forceinline__ void copy_words_home(const InstDesc &D) {
    const size_t nw = (size_t)D.tiles_x * (size_t)D.H;
    const uint64_t *__restrict__ src = D.label_bits;
    uint64_t *__restrict__ dst = D.label_host;
    size_t done = 0;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
        const uint4 *__restrict__ s4 = reinterpret_cast<const uint4 *>(src);
        uint4 *__restrict__ d4 = reinterpret_cast<uint4 *>(dst);
        const size_t n4 = nw / 2;
        size_t i = threadIdx.x;
        // Named values, not an array: a private array here was moved to LDS
        // (64 bytes a thread).
        for (; i + 3 * SNT < n4; i += 4 * SNT) {
            const uint4 v0 = s4[i], v1 = s4[i + SNT], v2 = s4[i + 2 * SNT], v3 = s4[i + 3 * SNT];
            d4[i] = v0;
            d4[i + SNT] = v1;
            d4[i + 2 * SNT] = v2;
            d4[i + 3 * SNT] = v3;
        }
        for (; i < n4; i += SNT) d4[i] = s4[i];
        done = 2 * n4;
    }
    for (size_t i = done + threadIdx.x; i < nw; i += SNT) dst[i] = src[i];
}

template <int SNT>
__global__ __launch_bounds__(SNT) void k_sched(KParams p, int parity) {
    const int s = blockIdx.x;
    if (s == 0 && threadIdx.x == 0) p.ctl->q[parity ^ 1] = QCtl{0, 0, 0, 0};
    InstState &st = p.state[s];
    // The descriptor and the whole state are loaded up front, in one round
    // trip. This workgroup is the only writer of the state, so its counters
    // are written as loaded value + change, with no read-modify-write (a
    // dependent round trip that every wave would wait for at the next
    // barrier).
    const InstDesc D = p.desc[s];
    const InstState S0 = st;
    const int phase = S0.phase;
    if (phase == PH_IDLE || phase == PH_DONE) return;
    const int err = S0.err, sweeps = S0.sweeps, cur_cap = S0.cap, rounds0 = S0.gr_rounds, cur_partial = S0.partial;
    int pr_iter = S0.pr_iter;
    __syncthreads();  // every thread has read the state before thread 0 writes it
    if (threadIdx.x == 0) st.steps = S0.steps + 1;
    if ((phase == PH_START && err) || phase == PH_FIN) {
        // flow = sum of pos - sum of max(excess, 0) over the interior, from
        // the per-tile sums of the build and the FIN sweep.
        __shared__ int64_t red[SNT / 64];
        int64_t v = 0;
        if (phase == PH_FIN)
            for (int lt = threadIdx.x; lt < D.ntiles; lt += SNT) {
                const size_t t = (size_t)(D.tile0 + lt);
                v += p.tsum[2 * t] - p.tsum[2 * t + 1];
            }
        // The labels are final (the FIN sweep ran in the last step): their
        // host copy goes out now, while other instances are still solved.
        if (phase == PH_FIN && D.label_host && D.label_bits) copy_words_home<SNT>(D);
        const int64_t tot = block_sum64<SNT>(v, red);
        if (threadIdx.x == 0) {
            if (phase == PH_FIN) st.flow = S0.sum_pos + tot;
            st.phase = PH_DONE;
            atomicSub(&p.ctl->n_live, 1);
        }
        return;
    }
    // Capped relabels pay when flow moves locally (few rounds); an instance
    // still unfinished after CAP_ROUNDS relabel rounds has flow travelling
    // far, where capped rounds multiply the steps: it relabels exactly from
    // then on.
    // The first round (the one the build runs) is capped at gr_cap, the
    // next CAP_ROUNDS - 1 at gr_cap_late (>= gr_cap): a short cap while most
    // flow still moves locally, where instances that finish early do, and a
    // longer one for those that go on. A capped round with a larger cap than
    // the last one recomputes every live block: nodes frozen at the old cap
    // may lie below the new one.
    const int cap = (p.variant && p.gr_cap > 0 && rounds0 < CAP_ROUNDS) ? (rounds0 ? p.gr_cap_late : p.gr_cap) : DINF;
    const bool grow = cap < DINF && cur_cap > 0 && cur_cap < cap;
    const bool reinit = p.variant != 0;
    // Exact rounds after the first recompute only the blocks with a
    // reachable node within the stale distance of a push since the last round
    // (partial rounds). Exact distances never decrease and the labelling
    // before the round is valid, so a block reset to the sinks and relaxed
    // from its neighbours' distances gets valid labels no lower than before,
    // and blocks outside the round keep theirs (a neighbour's new distances
    // are no lower than its old ones, so no relaxation can change them). The
    // active nodes all lie in the round (a tile with one was pushed, which
    // marked it stale). Labels outside the round may be lower than exact,
    // which costs pushes, not correctness; the last round is a full one.
    const bool part = reinit && p.gr_cap > 0 && rounds0 > CAP_ROUNDS && cap >= DINF;
    if (phase == PH_START) {
        // A warm start relabels every tile: an edit may have revived tiles
        // the kept flags call unreachable (edits keep the masks current).
        start_global_relabel<SNT>(p, D, st, S0, parity, reinit, cap, true);
        return;
    }
    bool converged = false;  // a distance round has just converged (PH_GR -> PH_PR in this step)
    if (phase == PH_GR) {
        const int n = emit_tiles<SNT>(p, D, EM_DIRTY, cur_cap == CAP_FLOOD ? 1 : 0, false, OP_GR_RELAX, parity);
        if (n > 0) {
            if (threadIdx.x == 0) {
                st.gr_tiles = S0.gr_tiles + n;
                if (cur_cap >= DINF || cur_cap == CAP_FLOOD) st.gr_exact_tiles = S0.gr_exact_tiles + n;
            }
            return;
        }
        // The BFS has converged: finish, or push from the active tiles. After
        // a distance round the first push colour is emitted straight away
        // (below); finding no active tile there is the finishing test.
        const bool flood = cur_cap == CAP_FLOOD;
        if (flood && !any_active_tile<SNT>(p, D)) {
            if (threadIdx.x == 0) st.phase = PH_FIN;
            // The reachability round wrote the labels of every block it
            // visited; FIN labels the blocks it skipped (unreachable, or
            // reaching a sink everywhere).
            const int nf = emit_tiles<SNT>(p, D, EM_DEAD, 0, false, OP_FIN, parity);
            if (threadIdx.x == 0) st.fin_items = S0.fin_items + nf;
            return;
        }
        if (flood) {
            // An excess node can still reach a sink. Flood values are not
            // valid heights: relabel exactly before pushing (a capped round
            // could freeze the same nodes again).
            start_global_relabel<SNT>(p, D, st, S0, parity, reinit, DINF);
            return;
        }
        if (threadIdx.x == 0) st.phase = PH_PR;
        pr_iter = 0;
        converged = true;
    }
    // PH_PR: cycle through the tile colours, then relabel globally again.
    const int ncol = p.nbhd == 8 ? 4 : 2;
    if (pr_iter >= ncol * sweeps) {
        start_global_relabel<SNT>(p, D, st, S0, parity, reinit, cap, false, part, grow);
        return;
    }
    const int colour = pr_iter & (ncol - 1);
    int n = 0;
    for (int c = 0; c < ncol && n == 0; ++c)
        n = emit_tiles<SNT>(p, D, EM_COLOUR, (colour + c) & (ncol - 1), true, 0, parity);
    if (n == 0 && converged) {
        // A distance round converged with no active tile.
        if (cur_cap < DINF) {
            // Nothing active below the cap: check reachability before
            // finishing (a flood settles in far fewer tile visits than exact
            // distances; it also writes the labels).
            start_global_relabel<SNT>(p, D, st, S0, parity, reinit, CAP_FLOOD, false, false, false, cur_cap);
            return;
        }
        if (cur_partial) {
            // A partial round's labels outside its blocks are valid but may
            // be stale: a full exact round writes the labels (it finds
            // nothing active either, every excess node being at DINF).
            start_global_relabel<SNT>(p, D, st, S0, parity, reinit, DINF);
            return;
        }
        if (threadIdx.x == 0) st.phase = PH_FIN;
        // Wavefront kernels: the exact round wrote the labels of every block
        // it visited; FIN labels the blocks it skipped (unreachable).
        const int nf = emit_tiles<SNT>(p, D, p.variant ? EM_DEAD : EM_ALL, 0, false, OP_FIN, parity);
        if (threadIdx.x == 0) st.fin_items = S0.fin_items + nf;
        return;
    }
    if (n == 0) {
        start_global_relabel<SNT>(p, D, st, S0, parity, reinit, cap, false, part, grow);
        return;
    }
    if (threadIdx.x == 0) {
        st.pr_iter = pr_iter + 1;
        st.pr_tiles = S0.pr_tiles + n;
    }
    // The last colour pass before an exact round: that round's REINIT sweep
    // is queued now and runs right after this push pass, in the same step
    // (the relabel launch follows the push launch). Its tiles are every block
    // the previous round found a reachable node in (`live`, which push passes
    // do not write), and it reads the masks this pass leaves, so it is the
    // sweep the next step would have queued, one step sooner. A capped round
    // cannot start early: its tiles depend on the stale flags this pass sets.
    // Only after an exact round (cur_cap DINF): the push pass reads the
    // instance's cap, which starting the round sets, so it keeps the cap of
    // the round it follows.
    if (reinit && cur_cap >= DINF && cap >= DINF && !part && pr_iter + 1 >= ncol * sweeps)
        start_global_relabel<SNT>(p, D, st, S0, parity, reinit, cap);
}

// ---------------------------------------------------------------------------
// Tile push-relabel pass.
__global__ __launch_bounds__(NT) void k_pr(KParams p, int parity) {
    __shared__ int Hs[2][LH][LW];
    __shared__ int FLs[LH][LW], FRs[LH][LW], FUs[LH][LW], FDs[LH][LW];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Workgroups past the queue length have nothing to take: leave at once.
    const int qlen = p.ctl->q[parity].qlen_pr;
    if ((int)blockIdx.x >= qlen) return;
    for (int i = tid; i < LH * LW; i += NT) {
        (&FLs[0][0])[i] = 0;
        (&FRs[0][0])[i] = 0;
        (&FUs[0][0])[i] = 0;
        (&FDs[0][0])[i] = 0;
    }
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(&p.ctl->q[parity].head_pr, 1);
            s_item = q < qlen ? p.queue_pr[q] : -1;
        }
        __syncthreads();
        const int t = s_item;
        if (t < 0) break;
        const TileInfo ti = p.tiles[t];
        const int dthr = p.desc[ti.inst].n_interior;
        const size_t base = (size_t)t * TN;
        const int lx = lane + 1;
        int4 c[RPT];
        int e[RPT], d[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            c[k] = p.caps[g];
            e[k] = p.exc[g];
            d[k] = p.dist[g];
            Hs[0][y + 1][lx] = d[k];
            Hs[1][y + 1][lx] = d[k];
        }
        if (tid < TH) {
            const int y = tid;
            const int v = ti.nbr[0] >= 0 ? p.dist[(size_t)ti.nbr[0] * TN + y * TW + TW - 1] : DINF;
            Hs[0][y + 1][0] = v;
            Hs[1][y + 1][0] = v;
        } else if (tid < 2 * TH) {
            const int y = tid - TH;
            const int v = ti.nbr[1] >= 0 ? p.dist[(size_t)ti.nbr[1] * TN + y * TW] : DINF;
            Hs[0][y + 1][TW + 1] = v;
            Hs[1][y + 1][TW + 1] = v;
        }
        if (wave == 1) {
            const int v = ti.nbr[2] >= 0 ? p.dist[(size_t)ti.nbr[2] * TN + (TH - 1) * TW + lane] : DINF;
            Hs[0][0][lx] = v;
            Hs[1][0][lx] = v;
        } else if (wave == 2) {
            const int v = ti.nbr[3] >= 0 ? p.dist[(size_t)ti.nbr[3] * TN + lane] : DINF;
            Hs[0][TH + 1][lx] = v;
            Hs[1][TH + 1][lx] = v;
        }
        __syncthreads();

        int accL[RPT], accR[RPT];
        int accU = 0, accD = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) accL[k] = accR[k] = 0;
        int cur = 0, it = 0;
        while (it < p.inner_iters) {
            ++it;
            // Push along admissible arcs (d(u) == d(v) + 1) with the excess held
            // at the start of the iteration, in the fixed order l, r, u, d.
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int ly = wave * RPT + k + 1;
                int fl = 0, fr = 0, fu = 0, fd = 0;
                int ek = e[k];
                if (ek > 0 && d[k] < DINF) {
                    const int hd = d[k] - 1;
                    int4 ck = c[k];
                    if (ck.x > 0 && Hs[cur][ly][lx - 1] == hd) {
                        fl = min(ek, ck.x);
                        ek -= fl;
                        ck.x -= fl;
                    }
                    if (ek > 0 && ck.y > 0 && Hs[cur][ly][lx + 1] == hd) {
                        fr = min(ek, ck.y);
                        ek -= fr;
                        ck.y -= fr;
                    }
                    if (ek > 0 && ck.z > 0 && Hs[cur][ly - 1][lx] == hd) {
                        fu = min(ek, ck.z);
                        ek -= fu;
                        ck.z -= fu;
                    }
                    if (ek > 0 && ck.w > 0 && Hs[cur][ly + 1][lx] == hd) {
                        fd = min(ek, ck.w);
                        ek -= fd;
                        ck.w -= fd;
                    }
                    c[k] = ck;
                    e[k] = ek;
                }
                FLs[ly][lx] = fl;
                FRs[ly][lx] = fr;
                FUs[ly][lx] = fu;
                FDs[ly][lx] = fd;
                if (lane == 0) accL[k] += fl;
                if (lane == TW - 1) accR[k] += fr;
                if (ly == 1) accU += fu;
                if (ly == TH) accD += fd;
            }
            __syncthreads();
            // Receive, then relabel nodes left with excess and no admissible arc:
            // d = 1 + min{d(v) : r(u,v) > 0} over the heights of this iteration.
            int act = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int ly = wave * RPT + k + 1;
                const int inL = FRs[ly][lx - 1], inR = FLs[ly][lx + 1];
                const int inU = FDs[ly - 1][lx], inD = FUs[ly + 1][lx];
                e[k] += inL + inR + inU + inD;
                c[k].x += inL;
                c[k].y += inR;
                c[k].z += inU;
                c[k].w += inD;
                if (e[k] > 0 && d[k] < DINF) {
                    int m = DINF;
                    if (c[k].x > 0) m = min(m, Hs[cur][ly][lx - 1]);
                    if (c[k].y > 0) m = min(m, Hs[cur][ly][lx + 1]);
                    if (c[k].z > 0) m = min(m, Hs[cur][ly - 1][lx]);
                    if (c[k].w > 0) m = min(m, Hs[cur][ly + 1][lx]);
                    int nd = m >= DINF ? DINF : m + 1;
                    if (nd >= dthr) nd = DINF;
                    if (nd > d[k]) d[k] = nd;
                    act |= d[k] < DINF;
                }
                Hs[cur ^ 1][ly][lx] = d[k];
            }
            cur ^= 1;
            if (!__syncthreads_or(act)) break;
        }

        int mine = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            p.caps[g] = c[k];
            p.exc[g] = e[k];
            p.dist[g] = d[k];
            mine |= e[k] > 0 && d[k] < DINF;
        }
        // Flow pushed into halo nodes: excess and reverse residual of the
        // neighbour tile (not running in this colour pass; int32 atomics
        // because two same-colour tiles can feed one corner node).
        if (lane == 0 && ti.nbr[0] >= 0) {
            int any = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (accL[k]) {
                    const size_t gn = (size_t)ti.nbr[0] * TN + (wave * RPT + k) * TW + TW - 1;
                    atomicAdd(&p.exc[gn], accL[k]);
                    atomicAdd(reinterpret_cast<int *>(&p.caps[gn]) + 1, accL[k]);
                    any = 1;
                }
            if (any) set_active(p, ti.nbr[0], 1);
        }
        if (lane == TW - 1 && ti.nbr[1] >= 0) {
            int any = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (accR[k]) {
                    const size_t gn = (size_t)ti.nbr[1] * TN + (wave * RPT + k) * TW;
                    atomicAdd(&p.exc[gn], accR[k]);
                    atomicAdd(reinterpret_cast<int *>(&p.caps[gn]) + 0, accR[k]);
                    any = 1;
                }
            if (any) set_active(p, ti.nbr[1], 1);
        }
        if (wave == 0 && accU && ti.nbr[2] >= 0) {
            const size_t gn = (size_t)ti.nbr[2] * TN + (TH - 1) * TW + lane;
            atomicAdd(&p.exc[gn], accU);
            atomicAdd(reinterpret_cast<int *>(&p.caps[gn]) + 3, accU);
            set_active(p, ti.nbr[2], 1);
        }
        if (wave == NT / 64 - 1 && accD && ti.nbr[3] >= 0) {
            const size_t gn = (size_t)ti.nbr[3] * TN + lane;
            atomicAdd(&p.exc[gn], accD);
            atomicAdd(reinterpret_cast<int *>(&p.caps[gn]) + 2, accD);
            set_active(p, ti.nbr[3], 1);
        }
        const int any_active = __syncthreads_or(mine);
        if (tid == 0) {
            set_active(p, t, any_active);
            if (p.profile) atomic_add64(&p.ctl->pr_iters, it);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Global relabel (exact BFS by tile relaxation) and finalisation.
__global__ __launch_bounds__(NT) void k_gr(KParams p, int parity) {
    __shared__ int Ds[LH][LW];
    __shared__ int s_item;
    __shared__ int64_t red[NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lx = lane + 1;
    const int qlen = p.ctl->q[parity].qlen_gr;
    if ((int)blockIdx.x >= qlen) return;
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(&p.ctl->q[parity].head_gr, 1);
            s_item = q < qlen ? p.queue_gr[q] : -1;
        }
        __syncthreads();
        const int item = s_item;
        if (item < 0) break;
        const int op = item & 7, t = item >> 3;
        const TileInfo ti = p.tiles[t];
        const InstDesc D = p.desc[ti.inst];  // by value: stores below cannot force reloads
        const size_t base = (size_t)t * TN;

        if (op == OP_FIN) {
            const int64_t sx = fin_rows_wg(p, D, ti, base, wave, lane);
            const int64_t tot = block_sum64(sx, red);
            if (tid == 0) p.tsum[2 * (size_t)t + 1] = tot;
            __syncthreads();
            continue;
        }

        const bool init = op == OP_GR_INIT;
        int dv[RPT], orig[RPT], m[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            int mk, d0;
            if (init) {
                const int4 cc = p.caps[g];
                const int ee = p.exc[g];
                mk = (cc.x > 0) | ((cc.y > 0) << 1) | ((cc.z > 0) << 2) | ((cc.w > 0) << 3) | ((ee > 0) << 4) |
                     ((ee < 0) << 5);
                p.mask[g] = (uint8_t)mk;
                d0 = ee < 0 ? 0 : DINF;
            } else {
                mk = p.mask[g];
                d0 = p.dist[g];
            }
            m[k] = mk;
            dv[k] = orig[k] = d0;
            Ds[y + 1][lx] = d0;
        }
        // Halo: initial distances (from the neighbour's excess) for INIT,
        // current distances otherwise. Both are upper bounds of the BFS
        // distance, so stale reads only delay convergence.
        if (tid < TH) {
            const int y = tid;
            int v = DINF;
            if (ti.nbr[0] >= 0) {
                const size_t gn = (size_t)ti.nbr[0] * TN + y * TW + TW - 1;
                v = init ? (p.exc[gn] < 0 ? 0 : DINF) : p.dist[gn];
            }
            Ds[y + 1][0] = v;
        } else if (tid < 2 * TH) {
            const int y = tid - TH;
            int v = DINF;
            if (ti.nbr[1] >= 0) {
                const size_t gn = (size_t)ti.nbr[1] * TN + y * TW;
                v = init ? (p.exc[gn] < 0 ? 0 : DINF) : p.dist[gn];
            }
            Ds[y + 1][TW + 1] = v;
        }
        if (wave == 1) {
            int v = DINF;
            if (ti.nbr[2] >= 0) {
                const size_t gn = (size_t)ti.nbr[2] * TN + (TH - 1) * TW + lane;
                v = init ? (p.exc[gn] < 0 ? 0 : DINF) : p.dist[gn];
            }
            Ds[0][lx] = v;
        } else if (wave == 2) {
            int v = DINF;
            if (ti.nbr[3] >= 0) {
                const size_t gn = (size_t)ti.nbr[3] * TN + lane;
                v = init ? (p.exc[gn] < 0 ? 0 : DINF) : p.dist[gn];
            }
            Ds[TH + 1][lx] = v;
        }
        __syncthreads();

        bool conv = false;
        for (int it = 0; it < GR_MAX_ITERS; ++it) {
            int ch = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int mk = m[k];
                if ((mk & 15) && dv[k] > 0) {
                    const int ly = wave * RPT + k + 1;
                    int best = dv[k];
                    if (mk & 1) best = min(best, Ds[ly][lx - 1] + 1);
                    if (mk & 2) best = min(best, Ds[ly][lx + 1] + 1);
                    if (mk & 4) best = min(best, Ds[ly - 1][lx] + 1);
                    if (mk & 8) best = min(best, Ds[ly + 1][lx] + 1);
                    if (best < dv[k]) {
                        dv[k] = best;
                        Ds[ly][lx] = best;
                        ch = 1;
                    }
                }
            }
            if (!__syncthreads_or(ch)) {
                conv = true;
                break;
            }
        }

        int chL = 0, chR = 0, chU = 0, chD = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            const int changed = dv[k] != orig[k];
            if (changed || init) p.dist[g] = dv[k];
            if (lane == 0) chL |= changed;
            if (lane == TW - 1) chR |= changed;
            if (y == 0) chU |= changed;
            if (y == TH - 1) chD |= changed;
            cnt += (m[k] & 16) && dv[k] < DINF;
        }
        chL = __syncthreads_or(chL);
        chR = __syncthreads_or(chR);
        chU = __syncthreads_or(chU);
        chD = __syncthreads_or(chD);
        const int64_t total = block_sum64(cnt, red);
        if (tid == 0) {
            if (chL && ti.nbr[0] >= 0) mark_dirty(p, ti.nbr[0]);
            if (chR && ti.nbr[1] >= 0) mark_dirty(p, ti.nbr[1]);
            if (chU && ti.nbr[2] >= 0) mark_dirty(p, ti.nbr[2]);
            if (chD && ti.nbr[3] >= 0) mark_dirty(p, ti.nbr[3]);
            if (!conv) mark_dirty(p, t);
            set_active(p, t, total > 0);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 8-neighbourhood (NEIGHBORHOOD == 8, GraphCut.h:22,65-70): the workgroup
// kernels above with the four diagonal arcs (GraphCutKernels.cu:516-521 push
// them after the side arcs). Directions 0-7: left, right, up, down, up-left,
// up-right, down-left, down-right; OPP8 is the reverse arc. The LDS halo holds
// the four corner nodes too, tiles of one push pass never touch (4 colours),
// and flow into another tile's node is added there after the pass.
__device__ constexpr int DX8[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
__device__ constexpr int DY8[8] = {0, 0, -1, 1, -1, -1, 1, 1};
__device__ constexpr int OPP8[8] = {1, 0, 3, 2, 7, 6, 5, 4};

__device__ __forceinline__ int cap8(const int4 &c, const int4 &c8, int dir) {
    switch (dir) {
        case 0: return c.x;
        case 1: return c.y;
        case 2: return c.z;
        case 3: return c.w;
        case 4: return c8.x;
        case 5: return c8.y;
        case 6: return c8.z;
        default: return c8.w;
    }
}

__device__ __forceinline__ void add_cap8(int4 &c, int4 &c8, int dir, int v) {
    switch (dir) {
        case 0: c.x += v; break;
        case 1: c.y += v; break;
        case 2: c.z += v; break;
        case 3: c.w += v; break;
        case 4: c8.x += v; break;
        case 5: c8.y += v; break;
        case 6: c8.z += v; break;
        default: c8.w += v; break;
    }
}

// Tile id and in-tile index of the node at tile-local (x, y) one step outside
// the tile (or inside it: returns t).
__device__ __forceinline__ int tile_of8(const TileInfo &ti, const int4 &n8, int t, int x, int y) {
    const int sx = x < 0 ? -1 : (x >= TW ? 1 : 0), sy = y < 0 ? -1 : (y >= TH ? 1 : 0);
    if (sx == 0 && sy == 0) return t;
    if (sy == 0) return sx < 0 ? ti.nbr[0] : ti.nbr[1];
    if (sx == 0) return sy < 0 ? ti.nbr[2] : ti.nbr[3];
    if (sy < 0) return sx < 0 ? n8.x : n8.y;
    return sx < 0 ? n8.z : n8.w;
}

// Load the halo ring (sides and corners) of array `a` of tile t's neighbours
// into S[LH][LW] (and S2 if given); `init` reads the initial BFS distance (0
// where excess < 0) instead of the stored distance.
__device__ __forceinline__ void load_halo8(const KParams &p, const TileInfo &ti, const int4 &n8, int tid,
                                           int (*S)[LW], int (*S2)[LW], bool init) {
    // Ring positions: 2 * TW (top, bottom rows) + 2 * TH (left, right columns) + 4 corners.
    constexpr int RING = 2 * TW + 2 * TH + 4;
    for (int r = tid; r < RING; r += NT) {
        int lx, ly;
        if (r < TW) { lx = r + 1; ly = 0; }
        else if (r < 2 * TW) { lx = r - TW + 1; ly = TH + 1; }
        else if (r < 2 * TW + TH) { lx = 0; ly = r - 2 * TW + 1; }
        else if (r < 2 * TW + 2 * TH) { lx = TW + 1; ly = r - 2 * TW - TH + 1; }
        else { const int c = r - 2 * TW - 2 * TH; lx = (c & 1) ? TW + 1 : 0; ly = (c & 2) ? TH + 1 : 0; }
        const int x = lx - 1, y = ly - 1;
        const int tn = tile_of8(ti, n8, -1, x, y);
        int v = DINF;
        if (tn >= 0) {
            const size_t gn = (size_t)tn * TN + ((y + TH) % TH) * TW + (x + TW) % TW;
            v = init ? (p.exc[gn] < 0 ? 0 : DINF) : p.dist[gn];
        }
        S[ly][lx] = v;
        if (S2) S2[ly][lx] = v;
    }
}

__global__ __launch_bounds__(NT) void k_pr8(KParams p, int parity) {
    __shared__ int Hs[2][LH][LW];
    __shared__ int Fs[8][LH][LW];
    __shared__ int s_item;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int qlen = p.ctl->q[parity].qlen_pr;
    if ((int)blockIdx.x >= qlen) return;
    for (int i = tid; i < 8 * LH * LW; i += NT) (&Fs[0][0][0])[i] = 0;  // halo cells stay 0
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(&p.ctl->q[parity].head_pr, 1);
            s_item = q < qlen ? p.queue_pr[q] : -1;
        }
        __syncthreads();
        const int t = s_item;
        if (t < 0) break;
        const TileInfo ti = p.tiles[t];
        const int4 n8 = p.nbr8[t];
        const int dthr = p.desc[ti.inst].n_interior;
        const size_t base = (size_t)t * TN;
        const int lx = lane + 1;
        int4 c[RPT], c8[RPT];
        int e[RPT], d[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            c[k] = p.caps[g];
            c8[k] = p.caps8[g];
            e[k] = p.exc[g];
            d[k] = p.dist[g];
            Hs[0][y + 1][lx] = d[k];
            Hs[1][y + 1][lx] = d[k];
        }
        load_halo8(p, ti, n8, tid, Hs[0], Hs[1], false);
        __syncthreads();

        // Flow pushed out of the tile over the pass, per boundary node and arc.
        int acc[RPT][8];
#pragma unroll
        for (int k = 0; k < RPT; ++k)
#pragma unroll
            for (int dir = 0; dir < 8; ++dir) acc[k][dir] = 0;
        int cur = 0, it = 0;
        while (it < p.inner_iters) {
            ++it;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int ly = wave * RPT + k + 1;
                int ek = e[k];
                const bool can = ek > 0 && d[k] < DINF;
                const int hd = d[k] - 1;
#pragma unroll
                for (int dir = 0; dir < 8; ++dir) {
                    int f = 0;
                    if (can && ek > 0) {
                        const int cv = cap8(c[k], c8[k], dir);
                        if (cv > 0 && Hs[cur][ly + DY8[dir]][lx + DX8[dir]] == hd) {
                            f = min(ek, cv);
                            ek -= f;
                            add_cap8(c[k], c8[k], dir, -f);
                        }
                    }
                    Fs[dir][ly][lx] = f;
                    acc[k][dir] += f;
                }
                e[k] = ek;
            }
            __syncthreads();
            // Receive along every arc, then relabel nodes left with excess.
            int act = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int ly = wave * RPT + k + 1;
#pragma unroll
                for (int dir = 0; dir < 8; ++dir) {
                    const int in = Fs[OPP8[dir]][ly + DY8[dir]][lx + DX8[dir]];
                    e[k] += in;
                    add_cap8(c[k], c8[k], dir, in);
                }
                if (e[k] > 0 && d[k] < DINF) {
                    int m = DINF;
#pragma unroll
                    for (int dir = 0; dir < 8; ++dir)
                        if (cap8(c[k], c8[k], dir) > 0) m = min(m, Hs[cur][ly + DY8[dir]][lx + DX8[dir]]);
                    int nd = m >= DINF ? DINF : m + 1;
                    if (nd >= dthr) nd = DINF;
                    if (nd > d[k]) d[k] = nd;
                    act |= d[k] < DINF;
                }
                Hs[cur ^ 1][ly][lx] = d[k];
            }
            cur ^= 1;
            if (!__syncthreads_or(act)) break;
        }

        int mine = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            p.caps[g] = c[k];
            p.caps8[g] = c8[k];
            p.exc[g] = e[k];
            p.dist[g] = d[k];
            mine |= e[k] > 0 && d[k] < DINF;
        }
        // Flow into nodes of other tiles (none of which runs in this pass):
        // their excess and the reverse residual arc, by int32 atomics (two
        // tiles of one colour can feed the same node).
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
#pragma unroll
            for (int dir = 0; dir < 8; ++dir) {
                const int a = acc[k][dir];
                const int x2 = lane + DX8[dir], y2 = y + DY8[dir];
                if (a == 0 || (x2 >= 0 && x2 < TW && y2 >= 0 && y2 < TH)) continue;
                const int tn = tile_of8(ti, n8, t, x2, y2);
                if (tn < 0) continue;
                const size_t gn = (size_t)tn * TN + ((y2 + TH) % TH) * TW + (x2 + TW) % TW;
                atomicAdd(&p.exc[gn], a);
                const int o = OPP8[dir];
                if (o < 4)
                    atomicAdd(reinterpret_cast<int *>(&p.caps[gn]) + o, a);
                else
                    atomicAdd(reinterpret_cast<int *>(&p.caps8[gn]) + (o - 4), a);
                set_active(p, tn, 1);
            }
        }
        const int any_active = __syncthreads_or(mine);
        if (tid == 0) {
            set_active(p, t, any_active);
            if (p.profile) atomic_add64(&p.ctl->pr_iters, it);
        }
        __syncthreads();
    }
}

// Global relabel and finalisation on the 8-neighbourhood (as k_gr).
__global__ __launch_bounds__(NT) void k_gr8(KParams p, int parity) {
    __shared__ int Ds[LH][LW];
    __shared__ int s_item;
    __shared__ int64_t red[NT / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lx = lane + 1;
    const int qlen = p.ctl->q[parity].qlen_gr;
    if ((int)blockIdx.x >= qlen) return;
    for (;;) {
        if (tid == 0) {
            const int q = atomicAdd(&p.ctl->q[parity].head_gr, 1);
            s_item = q < qlen ? p.queue_gr[q] : -1;
        }
        __syncthreads();
        const int item = s_item;
        if (item < 0) break;
        const int op = item & 7, t = item >> 3;
        const TileInfo ti = p.tiles[t];
        const int4 n8 = p.nbr8[t];
        const InstDesc D = p.desc[ti.inst];
        const size_t base = (size_t)t * TN;

        if (op == OP_FIN) {
            const int64_t sx = fin_rows_wg(p, D, ti, base, wave, lane);
            const int64_t tot = block_sum64(sx, red);
            if (tid == 0) p.tsum[2 * (size_t)t + 1] = tot;
            __syncthreads();
            continue;
        }

        const bool init = op == OP_GR_INIT;
        int dv[RPT], orig[RPT], m[RPT];
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            int mk, d0;
            if (init) {
                const int4 cc = p.caps[g], c8 = p.caps8[g];
                const int ee = p.exc[g];
                mk = (cc.x > 0) | ((cc.y > 0) << 1) | ((cc.z > 0) << 2) | ((cc.w > 0) << 3) | ((ee > 0) << 4) |
                     ((ee < 0) << 5);
                const int m8 = (c8.x > 0) | ((c8.y > 0) << 1) | ((c8.z > 0) << 2) | ((c8.w > 0) << 3);
                p.mask[g] = (uint8_t)mk;
                p.mask8[g] = (uint8_t)m8;
                mk |= m8 << 8;
                d0 = ee < 0 ? 0 : DINF;
            } else {
                mk = p.mask[g] | (p.mask8[g] << 8);
                d0 = p.dist[g];
            }
            m[k] = mk;
            dv[k] = orig[k] = d0;
            Ds[y + 1][lx] = d0;
        }
        // Halo (sides and corners): upper bounds of the BFS distance, so stale
        // reads only delay convergence.
        load_halo8(p, ti, n8, tid, Ds, nullptr, init);
        __syncthreads();

        bool conv = false;
        for (int it = 0; it < GR_MAX_ITERS; ++it) {
            int ch = 0;
#pragma unroll
            for (int k = 0; k < RPT; ++k) {
                const int mk = m[k];
                if ((mk & 0xF0F) && dv[k] > 0) {
                    const int ly = wave * RPT + k + 1;
                    int best = dv[k];
#pragma unroll
                    for (int dir = 0; dir < 8; ++dir) {
                        const int bit = dir < 4 ? (1 << dir) : (1 << (dir + 4));
                        if (mk & bit) best = min(best, Ds[ly + DY8[dir]][lx + DX8[dir]] + 1);
                    }
                    if (best < dv[k]) {
                        dv[k] = best;
                        Ds[ly][lx] = best;
                        ch = 1;
                    }
                }
            }
            if (!__syncthreads_or(ch)) {
                conv = true;
                break;
            }
        }

        int chL = 0, chR = 0, chU = 0, chD = 0, cnt = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int y = wave * RPT + k;
            const size_t g = base + y * TW + lane;
            const int changed = dv[k] != orig[k];
            if (changed || init) p.dist[g] = dv[k];
            if (lane == 0) chL |= changed;
            if (lane == TW - 1) chR |= changed;
            if (y == 0) chU |= changed;
            if (y == TH - 1) chD |= changed;
            // A changed corner node is in the halo of the diagonal tile.
            if (changed && (lane == 0 || lane == TW - 1) && (y == 0 || y == TH - 1)) {
                const int tn = y == 0 ? (lane == 0 ? n8.x : n8.y) : (lane == 0 ? n8.z : n8.w);
                if (tn >= 0) mark_dirty(p, tn);
            }
            cnt += (m[k] & 16) && dv[k] < DINF;
        }
        chL = __syncthreads_or(chL);
        chR = __syncthreads_or(chR);
        chU = __syncthreads_or(chU);
        chD = __syncthreads_or(chD);
        const int64_t total = block_sum64(cnt, red);
        if (tid == 0) {
            if (chL && ti.nbr[0] >= 0) mark_dirty(p, ti.nbr[0]);
            if (chR && ti.nbr[1] >= 0) mark_dirty(p, ti.nbr[1]);
            if (chU && ti.nbr[2] >= 0) mark_dirty(p, ti.nbr[2]);
            if (chD && ti.nbr[3] >= 0) mark_dirty(p, ti.nbr[3]);
            if (!conv) mark_dirty(p, t);
            set_active(p, t, total > 0);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Wavefront-per-tile kernels. Lane x of a 64-lane workgroup owns column x of
// the tile: rows 0..15 live in registers, so vertical neighbours are register
// reads and horizontal neighbours are one cross-lane move (DPP wave shift for
// distance-1, ds_bpermute shuffles for the scan). No LDS arrays, no workgroup
// barriers, no per-tile atomics; flags are plain stores read by the scheduler.

// lane x <- lane x-1 (lane 0 gets 0) / lane x <- lane x+1 (lane 63 gets 0).
__device__ __forceinline__ int shr1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xF, 0xF, false); }
__device__ __forceinline__ int shl1(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xF, 0xF, false); }

// 64-bit wavefront sum with DPP moves only (no LDS permutes): Hillis-Steele
// within 16-lane rows, then row_bcast:15 / row_bcast:31; lane 63 ends with
// the total, which every lane reads back.
__device__ __forceinline__ int64_t wave_sum64_dpp(int64_t v) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)((uint64_t)v >> 32);
    auto step = [&](auto ctrl_tag, int rows) {
        constexpr int ctrl = decltype(ctrl_tag)::value;
        uint32_t plo, phi;
        if (rows == 0xF) {
            plo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, 0xF, 0xF, false);
            phi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, 0xF, 0xF, false);
        } else if (rows == 0xA) {
            plo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, 0xA, 0xF, false);
            phi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, 0xA, 0xF, false);
        } else {
            plo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)lo, ctrl, 0xC, 0xF, false);
            phi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)hi, ctrl, 0xC, 0xF, false);
        }
        const uint64_t sum = ((uint64_t)hi << 32 | lo) + ((uint64_t)phi << 32 | plo);
        lo = (uint32_t)sum;
        hi = (uint32_t)(sum >> 32);
    };
    step(std::integral_constant<int, 0x111>{}, 0xF);
    step(std::integral_constant<int, 0x112>{}, 0xF);
    step(std::integral_constant<int, 0x114>{}, 0xF);
    step(std::integral_constant<int, 0x118>{}, 0xF);
    step(std::integral_constant<int, 0x142>{}, 0xA);
    step(std::integral_constant<int, 0x143>{}, 0xC);
    const uint32_t tlo = (uint32_t)__builtin_amdgcn_readlane((int)lo, 63);
    const uint32_t thi = (uint32_t)__builtin_amdgcn_readlane((int)hi, 63);
    return (int64_t)((uint64_t)thi << 32 | tlo);
}

__device__ __forceinline__ int64_t wave_sum64_w(int64_t v) { return wave_sum64_dpp(v); }

// Halo distances: rows of the tiles above/below (one per lane) and the
// columns of the tiles left/right (lanes 0..15 hold the left column, lanes
// 16..31 the right column, indexed by row).
// Halo distances: rows of the tiles above/below (one per lane) and the
// columns of the tiles left/right (lanes 0..15 hold the left column, lanes
// 16..31 the right column, indexed by row). The loads are issued
// unconditionally from clamped addresses (a missing neighbour reads the own
// tile) and masked afterwards, so they overlap with the tile's own loads.
// HALO_WITNESS (reachability rounds): 0 where the stored distance is below the
// cap of the capped rounds (a residual path to a sink that is still intact,
// see k_grb), DINF elsewhere.
enum HaloSrc : int { HALO_DIST = 0, HALO_EXCESS = 1, HALO_MASK = 2, HALO_WITNESS = 3 };

// 8-neighbourhood: lanes 32..35 of the side halo register hold the four corner
// nodes, up-left (node (TW-1, TH-1) of the up-left tile), up-right ((0, TH-1)
// of the up-right tile), down-left ((TW-1, 0)) and down-right ((0, 0)).
constexpr int HS_UL = 2 * TH, HS_UR = 2 * TH + 1, HS_DL = 2 * TH + 2, HS_DR = 2 * TH + 3;

// Diagonal neighbour tiles from the side ones: an instance's tiles are
// row-major and contiguous, so the up-left tile is the up tile's left
// neighbour, and it exists when both the up and the left tile do.
__device__ __forceinline__ int tile_ul(const TileInfo &ti) { return ti.nbr[2] >= 0 && ti.nbr[0] >= 0 ? ti.nbr[2] - 1 : -1; }
__device__ __forceinline__ int tile_ur(const TileInfo &ti) { return ti.nbr[2] >= 0 && ti.nbr[1] >= 0 ? ti.nbr[2] + 1 : -1; }
__device__ __forceinline__ int tile_dl(const TileInfo &ti) { return ti.nbr[3] >= 0 && ti.nbr[0] >= 0 ? ti.nbr[3] - 1 : -1; }
__device__ __forceinline__ int tile_dr(const TileInfo &ti) { return ti.nbr[3] >= 0 && ti.nbr[1] >= 0 ? ti.nbr[3] + 1 : -1; }

// Side-halo node of a lane: its tile (-1: none) and in-tile index. Lanes
// 0..15 the left tile's column TW-1, 16..31 the right tile's column 0 (by
// row), and for the 8-neighbourhood lanes 32..35 the corner nodes.
template <int NB>
__device__ __forceinline__ int side_halo(const TileInfo &ti, int lane, int &idx) {
    if (lane < 2 * TH) {
        idx = (lane & (TH - 1)) * TW + (lane < TH ? TW - 1 : 0);
        return lane < TH ? ti.nbr[0] : ti.nbr[1];
    }
    if (NB == 8 && lane <= HS_DR) {
        idx = (lane == HS_UL ? (TH - 1) * TW + TW - 1 : lane == HS_UR ? (TH - 1) * TW : lane == HS_DL ? TW - 1 : 0);
        return lane == HS_UL ? tile_ul(ti) : lane == HS_UR ? tile_ur(ti) : lane == HS_DL ? tile_dl(ti) : tile_dr(ti);
    }
    idx = 0;
    return -1;
}

template <int SRC, int NB = 4>
__device__ __forceinline__ void load_halo(const KParams &p, const TileInfo &ti, int t, int lane, int &hU, int &hD,
                                          int &hs, int wk = 0) {
    const int nu = ti.nbr[2], nd = ti.nbr[3];
    int sidx;
    const int side = side_halo<NB>(ti, lane, sidx);
    const size_t gu = (size_t)(nu >= 0 ? nu : t) * TN + (size_t)(TH - 1) * TW + lane;
    const size_t gd = (size_t)(nd >= 0 ? nd : t) * TN + lane;
    const size_t gs = (size_t)(side >= 0 ? side : t) * TN + (size_t)sidx;
    int vu, vd, vs;
    if (SRC == HALO_EXCESS) {
        vu = p.exc[gu] < 0 ? 0 : DINF;
        vd = p.exc[gd] < 0 ? 0 : DINF;
        vs = p.exc[gs] < 0 ? 0 : DINF;
    } else if (SRC == HALO_MASK) {
        vu = (p.mask[gu] & MB_ENEG) ? 0 : DINF;
        vd = (p.mask[gd] & MB_ENEG) ? 0 : DINF;
        vs = (p.mask[gs] & MB_ENEG) ? 0 : DINF;
    } else if (SRC == HALO_WITNESS) {
        vu = p.dist[gu] < wk ? 0 : DINF;
        vd = p.dist[gd] < wk ? 0 : DINF;
        vs = p.dist[gs] < wk ? 0 : DINF;
    } else {
        vu = p.dist[gu];
        vd = p.dist[gd];
        vs = p.dist[gs];
    }
    hU = nu >= 0 ? vu : DINF;
    hD = nd >= 0 ? vd : DINF;
    hs = side >= 0 ? vs : DINF;
}

__device__ __forceinline__ int dmin_sat(int a, int b) { return min(min(a, b), DINF); }

__device__ __forceinline__ int node_mask(const int4 &c, int e) {
    return (c.x > 0) | ((c.y > 0) << 1) | ((c.z > 0) << 2) | ((c.w > 0) << 3) | ((e > 0) ? MB_EPOS : 0) |
           ((e < 0) ? MB_ENEG : 0);
}

// Per-tile arrays through a wave-uniform base pointer: row y of lane x is
// element y * TW + x, so loads and stores take a scalar base, one 32-bit
// lane offset and immediate row offsets instead of a 64-bit address per row.
struct TileRef {
    int4 *c;
    int32_t *e, *d;
    uint8_t *m;
};
__device__ __forceinline__ TileRef tile_ref(const KParams &p, int t) {
    const size_t o = (size_t)t * TN;
    return TileRef{p.caps + o, p.exc + o, p.dist + o, p.mask + o};
}

// Flow f > 0 was pushed into node g of a tile that is not running: its excess
// became e_new and the residual arc back to the pusher (bit arc) is open.
// Within one colour pass such a node only gains excess, so setting "excess
// > 0" and clearing "excess < 0" are monotone and commute between pushers.
__device__ __forceinline__ void mask_after_push_in(const KParams &p, size_t g, int arc, int e_new) {
    uint32_t *w = reinterpret_cast<uint32_t *>(p.mask) + (g >> 2);
    const int sh = (int)(g & 3) * 8;
    atomicOr(w, (uint32_t)(arc | (e_new > 0 ? MB_EPOS : 0)) << sh);
    if (e_new >= 0) atomicAnd(w, ~((uint32_t)MB_ENEG << sh));
}

// Diagonal arcs back to the pushers opened on a node of a tile that is not
// running (8-neighbourhood mask8 bits: 1 up-left, 2 up-right, 4 down-left,
// 8 down-right).
__device__ __forceinline__ void mask8_after_push_in(const KParams &p, size_t g, int arcs) {
    uint32_t *w = reinterpret_cast<uint32_t *>(p.mask8) + (g >> 2);
    atomicOr(w, (uint32_t)arcs << ((int)(g & 3) * 8));
}

// (min, +) composition scans along a 64-lane row. Lane x holds the map
// f_x(v) = min(A, v + B); the scans leave in lane x the composition of the
// maps of every lane between x and the row end the values come from (lane 0
// for scan_from_left, lane 63 for scan_from_right), farthest applied first.
// Composing g after f: (A, B) = (min(A_g, A_f + B_g), min(B_f + B_g, INF)).
// Identity: (INF, 0).
__device__ __forceinline__ void compose_after(int &A, int &B, int pA, int pB) {
    A = min(A, pA + B);
    B = min(pB + B, DINF);
}

// The composition steps carry Ac = INF - A, whose identity is 0: a lane
// without a source in the DPP pattern reads 0, so the shifted value and the
// subtraction fuse into one DPP instruction (with A itself every step first
// re-materialised INF as the move's old value). min(A, pA + B) becomes
// max(Ac, pAc - B).
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_compose_c(int &Ac, int &B) {
    const int pAc = __builtin_amdgcn_update_dpp(0, Ac, CTRL, ROWS, 0xF, false);
    const int pB = __builtin_amdgcn_update_dpp(0, B, CTRL, ROWS, 0xF, false);
    Ac = max(Ac, pAc - B);
    B = min(pB + B, DINF);
}

// From lane 0 upwards: Hillis-Steele within each 16-lane row (row_shr 1, 2,
// 4, 8), then row_bcast:15 (row 0 -> row 1, row 2 -> row 3) and
// row_bcast:31 (rows 0..1 -> rows 2 and 3).
__device__ __forceinline__ void scan_from_left(int &A, int &B) {
    int Ac = DINF - A;
    dpp_compose_c<0x111, 0xF>(Ac, B);
    dpp_compose_c<0x112, 0xF>(Ac, B);
    dpp_compose_c<0x114, 0xF>(Ac, B);
    dpp_compose_c<0x118, 0xF>(Ac, B);
    dpp_compose_c<0x142, 0xA>(Ac, B);
    dpp_compose_c<0x143, 0xC>(Ac, B);
    A = DINF - Ac;
}

// From lane 63 downwards: within each row by row_shl 1, 2, 4, 8; across rows
// from the row totals at lanes 16, 32 and 48 (scalar reads).
__device__ __forceinline__ void scan_from_right(int &A, int &B, int lane) {
    int Ac = DINF - A;
    dpp_compose_c<0x101, 0xF>(Ac, B);
    dpp_compose_c<0x102, 0xF>(Ac, B);
    dpp_compose_c<0x104, 0xF>(Ac, B);
    dpp_compose_c<0x108, 0xF>(Ac, B);
    A = DINF - Ac;
    const int a1 = __builtin_amdgcn_readlane(A, 16), b1 = __builtin_amdgcn_readlane(B, 16);
    const int a2 = __builtin_amdgcn_readlane(A, 32), b2 = __builtin_amdgcn_readlane(B, 32);
    const int a3 = __builtin_amdgcn_readlane(A, 48), b3 = __builtin_amdgcn_readlane(B, 48);
    // Suffixes over whole rows: S2 = row 3, S1 = row 3 then row 2, S0 = rows 3, 2, 1.
    int sa2 = a3, sb2 = b3;
    int sa1 = a2, sb1 = b2;
    compose_after(sa1, sb1, sa2, sb2);
    int sa0 = a1, sb0 = b1;
    compose_after(sa0, sb0, sa1, sb1);
    const int row = lane >> 4;
    const int pA = row == 0 ? sa0 : (row == 1 ? sa1 : (row == 2 ? sa2 : DINF));
    const int pB = row == 0 ? sb0 : (row == 1 ? sb1 : (row == 2 ? sb2 : 0));
    compose_after(A, B, pA, pB);
}

// Residual masks of a lane's 16 rows, packed four rows per register (bits
// 8k..8k+7 of w[y / 4] for row y = 4 * (y / 4) + k).
struct Masks {
    uint32_t w[TH / 4];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int i = 0; i < TH / 4; ++i) w[i] = 0;
    }
    __device__ __forceinline__ int get(int y) const { return (int)((w[y >> 2] >> ((y & 3) * 8)) & 0xffu); }
    __device__ __forceinline__ void put(int y, int v) { w[y >> 2] |= (uint32_t)(v & 0xff) << ((y & 3) * 8); }
};

// Relax d(u) = min(d(u), 1 + d(v)) over residual arcs u->v (mask bits: 1 left,
// 2 right, 4 up, 8 down) until no value changes. Per round: a downward and an
// upward sweep of each lane's column in registers, then for each row an exact
// propagation along the row in both directions by a (min, +) composition scan
// over the lanes -- lane x holds f_x(v) = min(A, v + B) (B = 1 if the arc to
// the neighbour exists, else DINF) and composing with the predecessors'
// prefix gives the row minimum including the halo value. A row is scanned
// only if a one-step check finds a value that can still decrease (otherwise
// the row is already at its fixed point). Returns true when converged; chg
// collects the rows whose value changed in this lane.
// 8-neighbourhood (m8: bits 1 up-left, 2 up-right, 4 down-left, 8 down-right):
// the downward sweep also relaxes the arcs into the row above's left and
// right neighbours, the upward sweep those into the row below's.
template <int NB = 4>
__device__ __forceinline__ bool relax_tile_wave(int (&d)[TH], Masks &m, Masks &m8, int hU, int hD, int hs, int lane,
                                                uint32_t &chg, int &rounds, int max_rounds) {
    rounds = 0;
    while (rounds < max_rounds) {
        ++rounds;
        // Keeps the per-row values derived from the masks and the side halo
        // (arc selects, scalar halo reads) inside the loop: hoisted, they
        // would hold about 32 vector and 64 scalar registers across it.
        asm volatile("" : "+v"(hs));
#pragma unroll
        for (int i = 0; i < TH / 4; ++i) asm volatile("" : "+v"(m.w[i]));
        if (NB == 8) {
#pragma unroll
            for (int i = 0; i < TH / 4; ++i) asm volatile("" : "+v"(m8.w[i]));
        }
        uint32_t rch = 0;
        // 4-neighbourhood sweeps as selects: written as nested ifs they
        // compiled to two exec-mask branches per row.
#pragma unroll
        for (int y = 0; y < TH; ++y) {
            const int U = y ? d[y - 1] : hU;
            if (NB == 4) {
                const int cand = (m.get(y) & 4) ? U + 1 : DINF;
                const bool lower = cand < d[y];
                rch |= lower ? 1u << y : 0u;
                d[y] = lower ? cand : d[y];
                // Row by row: scheduled freely, every row's candidate stayed
                // live until the change bits were gathered (43 VGPRs spilled).
                asm volatile("" : "+v"(d[y]), "+v"(rch));
                continue;
            }
            int cand = (m.get(y) & 4) ? U + 1 : DINF;
            {
                // lane x <- lane x - 1 / x + 1 of the row above; lanes 0 and 63
                // take the side halo's row y - 1 (the corners for row 0).
                const int sl = __builtin_amdgcn_readlane(hs, y ? y - 1 : HS_UL);
                const int sr = __builtin_amdgcn_readlane(hs, y ? TH + y - 1 : HS_UR);
                const int ul = __builtin_amdgcn_update_dpp(sl, U, 0x138, 0xF, 0xF, false);
                const int ur = __builtin_amdgcn_update_dpp(sr, U, 0x130, 0xF, 0xF, false);
                const int m8y = m8.get(y);
                if (m8y & 1) cand = min(cand, ul + 1);
                if (m8y & 2) cand = min(cand, ur + 1);
            }
            if (cand < d[y]) {
                d[y] = cand;
                rch |= 1u << y;
            }
        }
#pragma unroll
        for (int y = TH - 1; y >= 0; --y) {
            const int Dn = y < TH - 1 ? d[y + 1] : hD;
            if (NB == 4) {
                const int cand = (m.get(y) & 8) ? Dn + 1 : DINF;
                const bool lower = cand < d[y];
                rch |= lower ? 1u << y : 0u;
                d[y] = lower ? cand : d[y];
                // Row by row: scheduled freely, every row's candidate stayed
                // live until the change bits were gathered (43 VGPRs spilled).
                asm volatile("" : "+v"(d[y]), "+v"(rch));
                continue;
            }
            int cand = (m.get(y) & 8) ? Dn + 1 : DINF;
            {
                const int sl = __builtin_amdgcn_readlane(hs, y < TH - 1 ? y + 1 : HS_DL);
                const int sr = __builtin_amdgcn_readlane(hs, y < TH - 1 ? TH + y + 1 : HS_DR);
                const int dl = __builtin_amdgcn_update_dpp(sl, Dn, 0x138, 0xF, 0xF, false);
                const int dr = __builtin_amdgcn_update_dpp(sr, Dn, 0x130, 0xF, 0xF, false);
                const int m8y = m8.get(y);
                if (m8y & 4) cand = min(cand, dl + 1);
                if (m8y & 8) cand = min(cand, dr + 1);
            }
            if (cand < d[y]) {
                d[y] = cand;
                rch |= 1u << y;
            }
        }
        uint32_t rrow = 0;  // rows the row scans changed
#pragma unroll
        for (int y = 0; y < TH; ++y) {
            // Row y one lane over, lanes 0 and 63 taking the side halo's row
            // y (the DPP move's old value is kept where a lane has no source).
            const int hsl = __builtin_amdgcn_readlane(hs, y), hsr = __builtin_amdgcn_readlane(hs, TH + y);
            const int hl = __builtin_amdgcn_update_dpp(hsl, d[y], 0x138, 0xF, 0xF, false);
            const int hr = __builtin_amdgcn_update_dpp(hsr, d[y], 0x130, 0xF, 0xF, false);
            const int my = m.get(y);
            const bool can = ((my & 1) && hl + 1 < d[y]) || ((my & 2) && hr + 1 < d[y]);
            if (!__any(can)) continue;
            // left to right: arcs x -> x-1, halo = left column
            int A = d[y], B = (my & 1) ? 1 : DINF;
            scan_from_left(A, B);
            int nd = min(A, dmin_sat(hsl + B, DINF));
            // right to left: arcs x -> x+1, halo = right column
            int A2 = nd, B2 = (my & 2) ? 1 : DINF;
            scan_from_right(A2, B2, lane);
            nd = min(A2, dmin_sat(hsr + B2, DINF));
            const bool lower = nd < d[y];
            rrow |= lower ? 1u << y : 0u;
            d[y] = lower ? nd : d[y];
            asm volatile("" : "+v"(d[y]), "+v"(rrow));
        }
        rch |= rrow;
        chg |= rch;
        // 4-neighbourhood: a column's shortest paths are monotone, so the
        // downward and upward sweeps leave every column at its own fixed
        // point; if the row scans then change nothing, every row is at its
        // fixed point too and the tile has converged -- no confirming round.
        // (Diagonal arcs let a path go up and down again within the sweeps,
        // so the 8-neighbourhood waits for a round without any change.)
        if (!__any((NB == 4 ? rrow : rch) != 0)) return true;
    }
    return false;
}

// One bit per row of a lane's column (bit y = row y) for the mask bit `b` of
// every row: the four rows packed in a mask word sit at bits 0, 8, 16, 24;
// multiplying by 2^21 + 2^14 + 2^7 + 1 gathers them into bits 21..24
// without carries (all sixteen partial-product bits are distinct).
__device__ __forceinline__ uint32_t row_plane(const Masks &m, int b) {
    uint32_t plane = 0;
#pragma unroll
    for (int i = 0; i < TH / 4; ++i) {
        const uint32_t x = (m.w[i] >> b) & 0x01010101u;
        plane |= (((x * 0x00204081u) >> 21) & 0xfu) << (4 * i);
    }
    return plane;
}

// Diagonal arcs of a lane's 16 rows as row bit planes (8-neighbourhood).
struct DiagPlanes {
    uint32_t ul, ur, dl, dr;
};

template <int NB>
__device__ __forceinline__ DiagPlanes diag_planes(const Masks &m8) {
    if (NB != 8) return DiagPlanes{0, 0, 0, 0};
    return DiagPlanes{row_plane(m8, 0), row_plane(m8, 1), row_plane(m8, 2), row_plane(m8, 3)};
}

// One level-set step: S plus every node with a residual arc into S or into a
// halo node in the set. hv holds this lane's halo bits (bit 0 the node above
// row 0, bit 15 the node below row 15), lm / rm the side halo columns' bits
// (lane 0 takes lm as its left neighbour's set, lane 63 takes rm), and with
// diagonals cl / cr the corner bits lanes 0 / 63 take as the halo bits of
// the lane beyond the tile (bit 0 up-left / up-right, bit 15 down-left /
// down-right).
template <int NB>
__device__ __forceinline__ uint32_t level_step(uint32_t S, uint32_t PL, uint32_t PR, uint32_t PU, uint32_t PD,
                                               const DiagPlanes &P8, uint32_t hv, int lm, int rm, uint32_t cl,
                                               uint32_t cr) {
    // lane x <- lane x-1 (lane 0: the left halo column), lane x <- lane x+1 (lane 63: right halo)
    const uint32_t lf = (uint32_t)__builtin_amdgcn_update_dpp(lm, (int)S, 0x138, 0xF, 0xF, false);
    const uint32_t rt = (uint32_t)__builtin_amdgcn_update_dpp(rm, (int)S, 0x130, 0xF, 0xF, false);
    uint32_t Sn = S | (((S << 1) | (hv & 1u)) & PU) | (((S >> 1) | (hv & 0x8000u)) & PD) | (lf & PL) | (rt & PR);
    if (NB == 8) {
        const uint32_t hvl = (uint32_t)__builtin_amdgcn_update_dpp((int)cl, (int)hv, 0x138, 0xF, 0xF, false);
        const uint32_t hvr = (uint32_t)__builtin_amdgcn_update_dpp((int)cr, (int)hv, 0x130, 0xF, 0xF, false);
        Sn |= (((lf << 1) | (hvl & 1u)) & P8.ul) | (((lf >> 1) | (hvl & 0x8000u)) & P8.dl) |
              (((rt << 1) | (hvr & 1u)) & P8.ur) | (((rt >> 1) | (hvr & 0x8000u)) & P8.dr);
    }
    return Sn & 0xffffu;
}

// Halo bits of one level: nodes whose value passes `in` (<= k, or == 0).
template <int NB, typename In>
__device__ __forceinline__ void level_halo(int hU, int hD, int hs, int lane, In in, uint32_t &hv, int &lm, int &rm,
                                           uint32_t &cl, uint32_t &cr) {
    hv = (in(hU) ? 1u : 0u) | (in(hD) ? 0x8000u : 0u);
    const uint64_t bl = __ballot(lane < TH && in(hs));
    const uint64_t br = __ballot(lane >= TH && lane < 2 * TH && in(hs));
    lm = (int)((uint32_t)bl & 0xffffu);
    rm = (int)((uint32_t)(br >> TH) & 0xffffu);
    cl = cr = 0;
    if (NB == 8) {
        const uint64_t bc = __ballot(lane >= HS_UL && lane <= HS_DR && in(hs));
        cl = ((bc >> HS_UL) & 1u) | (((bc >> HS_DL) & 1u) << 15);
        cr = ((bc >> HS_UR) & 1u) | (((bc >> HS_DR) & 1u) << 15);
    }
}

// Capped BFS of one tile by level sets, 16 rows as bits of one register per
// lane: L_0 = sink nodes; L_(k+1) = L_k plus every node with a residual arc
// into L_k or into a halo node whose value is <= k. Then d = the first level
// holding the node, or K if none of L_0..L_(K-1) does; that is min(K, BFS
// distance) given the halo, the fixed point the relaxation reaches, in about
// 25 vector operations per level (4-neighbourhood). Level counts per row are
// kept bit-sliced (BITS bits, K < 2^BITS). chg collects rows whose value
// differs from d on entry.
template <int BITS, int NB = 4>
__device__ __forceinline__ void bfs_levels(int (&d)[TH], const Masks &m, const Masks &m8, int hU, int hD, int hs,
                                           int lane, int K, uint32_t &chg, int &levels) {
    const uint32_t PL = row_plane(m, 0), PR = row_plane(m, 1), PU = row_plane(m, 2), PD = row_plane(m, 3);
    const DiagPlanes P8 = diag_planes<NB>(m8);
    uint32_t S = row_plane(m, 5);
    uint32_t c[BITS];
#pragma unroll
    for (int b = 0; b < BITS; ++b) c[b] = 0;
    auto count = [&](uint32_t v) {
#pragma unroll
        for (int b = 0; b < BITS; ++b) {
            const uint32_t t = c[b] & v;
            c[b] ^= v;
            v = t;
        }
    };
    count(S);
    int j = 0;  // index of the last level computed
    for (int k = 0; k < K - 1; ++k) {
        uint32_t hv, cl, cr;
        int lm, rm;
        level_halo<NB>(hU, hD, hs, lane, [k](int v) { return v <= k; }, hv, lm, rm, cl, cr);
        S = level_step<NB>(S, PL, PR, PU, PD, P8, hv, lm, rm, cl, cr);
        count(S);
        j = k + 1;
    }
    levels = j + 1;
    uint32_t rch = 0;
#pragma unroll
    for (int y = 0; y < TH; ++y) {
        int cnt = 0;
#pragma unroll
        for (int b = 0; b < BITS; ++b) cnt |= (int)((c[b] >> y) & 1u) << b;
        const int nd = ((S >> y) & 1u) ? j + 1 - cnt : K;
        if (nd != d[y]) {
            d[y] = nd;
            rch |= 1u << y;
        }
    }
    chg |= rch;
}

// BFS of one tile whose every source is at distance 0 (the sinks and the
// halo nodes at 0: a REINIT sweep, whose halos come from the masks), by the
// same level sets run until they stop growing or reach the cap K (DINF: no
// cap): d = the first level holding the node, K for the nodes no level
// reaches. Exact for any K, in as many levels as the tile's longest BFS
// distance, where the relaxation (relax_tile_wave) needs a round per turn of
// a shortest path and more work per round. chg collects the rows whose value
// differs from d on entry.
template <int NB = 4>
__device__ __forceinline__ void bfs_levels_zero(int (&d)[TH], const Masks &m, const Masks &m8, int hU, int hD, int hs,
                                                int lane, int K, uint32_t &chg, int &levels) {
    const uint32_t PL = row_plane(m, 0), PR = row_plane(m, 1), PU = row_plane(m, 2), PD = row_plane(m, 3);
    const DiagPlanes P8 = diag_planes<NB>(m8);
    uint32_t S = row_plane(m, 5);
    uint32_t hv, cl, cr;
    int lm, rm;
    level_halo<NB>(hU, hD, hs, lane, [](int v) { return v == 0; }, hv, lm, rm, cl, cr);
    int nd[TH];
#pragma unroll
    for (int y = 0; y < TH; ++y) nd[y] = ((S >> y) & 1u) ? 0 : K;
    int L = 0;
    // S only grows and has at most TN bits; a level is computed only while L + 1 < K.
    while (L + 1 < K) {
        const uint32_t Sn = level_step<NB>(S, PL, PR, PU, PD, P8, hv, lm, rm, cl, cr);
        const uint32_t fresh = Sn & ~S;
        if (!__any(fresh != 0)) break;
        ++L;
#pragma unroll
        for (int y = 0; y < TH; ++y)
            if ((fresh >> y) & 1u) nd[y] = L;
        S = Sn;
    }
    levels = L + 1;
    uint32_t rch = 0;
#pragma unroll
    for (int y = 0; y < TH; ++y) {
        if (nd[y] != d[y]) {
            d[y] = nd[y];
            rch |= 1u << y;
        }
    }
    chg |= rch;
}

// Reachability flood of one tile by the same level sets, run to the fixed
// point: S = sinks, nodes already at 0 and nodes with a residual arc into S or
// into a halo node at 0; d = 0 on S, DINF elsewhere. Values only go from DINF
// to 0, so tiles re-queued by changed boundaries converge after few visits,
// where exact distances would be lowered again and again.
template <int NB = 4>
__device__ __forceinline__ void bfs_flood(int (&d)[TH], const Masks &m, const Masks &m8, int hU, int hD, int hs,
                                          int lane, uint32_t &chg, int &levels) {
    const uint32_t PL = row_plane(m, 0), PR = row_plane(m, 1), PU = row_plane(m, 2), PD = row_plane(m, 3);
    const DiagPlanes P8 = diag_planes<NB>(m8);
    uint32_t S = row_plane(m, 5);
#pragma unroll
    for (int y = 0; y < TH; ++y) S |= (d[y] == 0 ? 1u : 0u) << y;
    uint32_t hv, cl, cr;
    int lm, rm;
    level_halo<NB>(hU, hD, hs, lane, [](int v) { return v == 0; }, hv, lm, rm, cl, cr);
    int it = 0;
    for (;;) {  // S only grows and has at most TN bits: at most TN + 1 trips
        ++it;
        const uint32_t Sn = level_step<NB>(S, PL, PR, PU, PD, P8, hv, lm, rm, cl, cr);
        const bool grew = __any(Sn != S);
        S = Sn;
        if (!grew) break;
    }
    levels = it;
    uint32_t rch = 0;
#pragma unroll
    for (int y = 0; y < TH; ++y) {
        const int nd = ((S >> y) & 1u) ? 0 : DINF;
        if (nd != d[y]) {
            d[y] = nd;
            rch |= 1u << y;
        }
    }
    chg |= rch;
}

// Capped tile BFS for caps up to 15 by level sets; larger caps and exact
// rounds use the relaxation, reachability rounds the flood. Returns true when
// the tile converged.
template <int NB = 4>
__device__ __forceinline__ bool tile_bfs(int (&d)[TH], Masks &m, Masks &m8, int hU, int hD, int hs, int lane,
                                         int capv, uint32_t &chg, int &rounds, int max_rounds,
                                         bool zero_sources = false) {
    if (capv == CAP_FLOOD) {
        bfs_flood<NB>(d, m, m8, hU, hD, hs, lane, chg, rounds);
        return true;
    }
    if (capv <= 3) {
        bfs_levels<2, NB>(d, m, m8, hU, hD, hs, lane, capv, chg, rounds);
        return true;
    }
    if (capv <= 7) {
        bfs_levels<3, NB>(d, m, m8, hU, hD, hs, lane, capv, chg, rounds);
        return true;
    }
    if (capv <= 15) {
        bfs_levels<4, NB>(d, m, m8, hU, hD, hs, lane, capv, chg, rounds);
        return true;
    }
    if (zero_sources) {
        bfs_levels_zero<NB>(d, m, m8, hU, hD, hs, lane, capv, chg, rounds);
        return true;
    }
    return relax_tile_wave<NB>(d, m, m8, hU, hD, hs, lane, chg, rounds, max_rounds);
}

// Labels of one tile from converged exact distances: 1 for interior nodes
// that cannot reach the sink (SPREAD_ZEROS set, GraphCutKernels.cu:228-351),
// 0 elsewhere, border included.
__device__ __forceinline__ void write_labels(const InstDesc &D, int tx, int ty, int lane, const int (&d)[TH]) {
    const int X = tx * TW + lane;
    const bool colin = X > 0 && X < D.W - 1;
    uint64_t word = 0;  // lane y: the packed word of row y
#pragma unroll
    for (int y = 0; y < TH; ++y) {
        const int Y = ty * TH + y;
        const bool lab = colin && Y > 0 && Y < D.H - 1 && d[y] >= DINF;
        if (D.label_out && X < D.W && Y < D.H) stg_out(D.label_out, (size_t)Y * D.W + X, lab ? 1 : 0);
        if (D.label_bits) {
            const uint64_t b = __ballot(lab);
            if (lane == y) word = b;
        }
    }
    if (D.label_bits && lane < TH && ty * TH + lane < D.H)
        stg_out64(D.label_bits, (size_t)(ty * TH + lane) * D.tiles_x + tx, word);
}

// Anchor (top-left tile) of the global-relabel block holding tile (x, y).
__device__ __forceinline__ int block_anchor(const KParams &p, const InstDesc &D, int x, int y) {
    if (p.gbx == 1 && p.gby == 1) return D.tile0 + y * D.tiles_x + x;  // the common shape: no divisions
    return D.tile0 + (y - y % p.gby) * D.tiles_x + (x - x % p.gbx);
}

constexpr int GR_MAX_EXCHANGES = 64;

// Global relabel over a block of BX x BY tiles per workgroup, one wavefront
// per tile (lane x owns column x). Each wavefront relaxes its tile in
// registers; the block then swaps its inner boundary rows and columns
// through LDS and repeats until no inner halo value drops, so a BFS front
// crosses the whole block in one launch. Outer halos come from the
// neighbouring blocks' stored distances (upper bounds: a stale read only
// delays convergence); a block whose outer boundary changed marks its
// neighbours for the next step. FIN writes the labels of the block's tiles.
// The queue, tile table, descriptors and states are read-only in the
// wavefront kernels; passing them as separate restrict arguments lets their
// uniform loads go through the scalar cache instead of waiting behind the
// kernel's own vector stores.
#define GCFG_RO_ARGS                                                                                  \
    const int32_t *__restrict__ rq, const TileInfo *__restrict__ rtiles, const InstDesc *__restrict__ rdesc, \
        const InstState *__restrict__ rstate

// One-tile blocks fit 96 VGPRs without spilling: five wavefronts per SIMD
// instead of four (the bigger blocks would spill, so they keep the default;
// so would the 8-neighbourhood's one-tile blocks at five, they run four).
// One-tile blocks: GR_PACK independent wavefronts per workgroup, each a
// block of its own (no barrier joins them): the same wavefronts in a quarter
// of the workgroups. An empty kernel took 3.7 us to dispatch as 12 k one-wave
// workgroups and 1.6-1.8 us as 3 k of four waves (in a graph); in the solves
// the 4096^2 image went from 5.29 to 5.23 ms and 1024^2 from 0.182 to
// 0.178 ms (an empty launch still waits for its queue length from memory).
constexpr int GR_PACK = 4;
template <int BX, int BY>
constexpr int grb_threads() { return BX * BY == 1 ? 64 * GR_PACK : 64 * BX * BY; }

template <int BX, int BY, int NB = 4>
__global__ __launch_bounds__((grb_threads<BX, BY>())) __attribute__((amdgpu_waves_per_eu(BX * BY == 1 ? (NB == 8 ? 4 : 5) : 1))) void k_grb(
    KParams p, int parity, GCFG_RO_ARGS) {
    constexpr int NW = BX * BY;
    static_assert(NB == 4 || NW == 1, "8-neighbourhood relabels run one-tile blocks");
    __shared__ int s_top[NW][TW], s_bot[NW][TW], s_side[NW][2 * TH];
    const int qlen = p.ctl->q[parity].qlen_gr;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Block index and count: the workgroup's own for multi-tile blocks, the
    // wavefront's for packed one-tile blocks.
    const int blk = NW == 1 ? (int)blockIdx.x * GR_PACK + wave : (int)blockIdx.x;
    const int nblk = NW == 1 ? (int)gridDim.x * GR_PACK : (int)gridDim.x;
    if (blk >= qlen) return;
    const int w = NW == 1 ? 0 : wave;  // the tile's position in its block
    const bool leader = NW == 1 ? lane == 0 : threadIdx.x == 0;
    const int wi = w % BX, wj = w / BX;
    int64_t n_rounds = 0;
    // Each block takes a contiguous run of the queue: consecutive items
    // are usually neighbouring tiles of one instance, so their queue words,
    // tile records and descriptor hit the scalar cache, and their halo rows
    // are each other's tile rows. (Dealing positions out from a counter
    // instead balanced the runs but serialised thousands of same-address
    // atomics per launch: twice the time.)
    const int per = (qlen + nblk - 1) / nblk;
    const int q_end = min(qlen, (blk + 1) * per);
    for (int q = blk * per; q < q_end; ++q) {
        const int item = rq[q];
        const int op = item & 7, t0 = item >> 3;
        const int inst = rtiles[t0].inst;
        // By value: with a reference every label or flag store below could
        // alias the descriptor and force a reload before the next use.
        const InstDesc D = rdesc[inst];
        const int ax = rtiles[t0].tx, ay = rtiles[t0].ty;
        const int tx = ax + wi, ty = ay + wj;
        const bool have = tx < D.tiles_x && ty < D.tiles_y;
        const int t = have ? t0 + wj * D.tiles_x + wi : t0;
        const TileInfo ti = rtiles[t];
        const TileRef T = tile_ref(p, t);
        if (op == OP_FIN) {
            // Only blocks the last round skipped come here: every node
            // unreachable (labels 1) or every interior node reaching a sink
            // (labels 0); the others wrote their labels in that round.
            if (!have) continue;
            const int dv = p.live[t0] == 3 ? 0 : DINF;
            int dd[TH];
#pragma unroll
            for (int y = 0; y < TH; ++y) dd[y] = dv;
            write_labels(D, tx, ty, lane, dd);
            continue;
        }
        // (OP_GR_INIT is the workgroup kernels' op; here masks are rebuilt by
        // the push kernel's OP_MASK pass and every relabel starts with REINIT.)
        const bool reinit = op == OP_GR_REINIT;
        const int capv = rstate[inst].cap;  // DINF: exact, CAP_FLOOD: reachability, else the cap
        // A partial exact round resets only some blocks: their REINIT sweep
        // reads sink halos only, so each block queues itself for a sweep that
        // reads its neighbours' distances (those outside the round are final).
        const bool self_dirty = reinit && rstate[inst].partial;
        const int wcap = rstate[inst].wcap;  // a reachability round's witness threshold
        // "Near" (live 2: a node below the cap) is the current capped round's
        // cap; other rounds keep the option's (no capped round uses it after them).
        const int near_thr = (capv > 0 && capv < DINF) ? capv : p.gr_cap;
        const bool flood = capv == CAP_FLOOD;
        const int vinit = flood ? DINF : capv;  // start value of non-sink nodes
        int d[TH];
        Masks m, m8;
        m.clear();
        m8.clear();
        int hU = DINF, hD = DINF, hs = DINF;
        // A reachability round follows a capped round that converged with no
        // push pass in between, and skips the blocks at live 3 (every interior
        // node below the cap). A stored distance below the cap is then the
        // length of a residual path to a sink that still exists: the capped
        // round recomputed every block within the cap of a push since, and a
        // path shorter than the cap lies within the cap of its block. The
        // flood therefore starts from those nodes as well as from the sinks,
        // and reads its halos the same way, so a node whose only route to a
        // sink enters a skipped block through a node that is not itself a
        // sink still reaches it (the skipped block never marks anyone dirty).
        if (!have) {
#pragma unroll
            for (int y = 0; y < TH; ++y) d[y] = DINF;
        } else if (reinit && flood) {
            load_halo<HALO_WITNESS, NB>(p, ti, t, lane, hU, hD, hs, wcap);
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                m.put(y, T.m[y * TW + lane]);
                if (NB == 8) m8.put(y, p.mask8[(size_t)t * TN + y * TW + lane]);
                d[y] = T.d[y * TW + lane];
            }
#pragma unroll
            for (int y = 0; y < TH; ++y) d[y] = ((m.get(y) & MB_ENEG) || d[y] < wcap) ? 0 : DINF;
        } else if (reinit) {
            load_halo<HALO_MASK, NB>(p, ti, t, lane, hU, hD, hs);
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                m.put(y, T.m[y * TW + lane]);
                if (NB == 8) m8.put(y, p.mask8[(size_t)t * TN + y * TW + lane]);
            }
#pragma unroll
            for (int y = 0; y < TH; ++y) d[y] = (m.get(y) & MB_ENEG) ? 0 : vinit;
        } else if (flood) {
            // Visited blocks hold 0 or DINF by now; skipped ones their capped distances.
            load_halo<HALO_WITNESS, NB>(p, ti, t, lane, hU, hD, hs, wcap);
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                m.put(y, T.m[y * TW + lane]);
                if (NB == 8) m8.put(y, p.mask8[(size_t)t * TN + y * TW + lane]);
                d[y] = T.d[y * TW + lane];
            }
        } else {
            load_halo<HALO_DIST, NB>(p, ti, t, lane, hU, hD, hs);
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                m.put(y, T.m[y * TW + lane]);
                if (NB == 8) m8.put(y, p.mask8[(size_t)t * TN + y * TW + lane]);
                d[y] = T.d[y * TW + lane];
            }
        }
        // Inner neighbours inside the block (the tile exists and is in the block).
        const bool inU = wj > 0, inD = wj < BY - 1 && ty + 1 < D.tiles_y;
        const bool inL = wi > 0, inR = wi < BX - 1 && tx + 1 < D.tiles_x;
        uint32_t chg = 0;
        bool capped = false;
        bool need = have;  // this tile has a new halo value or did not converge yet
        for (int xr = 0;; ++xr) {
            int rounds = 0;
            bool conv = true;
            // The first sweep of a REINIT starts from sinks and mask halos only (every
            // source at 0); later exchanges bring block-inner halos of any value.
            if (need)
                conv = tile_bfs<NB>(d, m, m8, hU, hD, hs, lane, capv, chg, rounds, p.gr_rounds, reinit && !flood && xr == 0);
            n_rounds += rounds;
            if constexpr (NW == 1) {
                capped = !conv;
                break;
            } else {
                s_top[w][lane] = d[0];
                s_bot[w][lane] = d[TH - 1];
                if (lane == 0 || lane == TW - 1) {
                    const int o = lane ? TH : 0;  // column 0 at [0, 16), column 63 at [16, 32)
#pragma unroll
                    for (int y = 0; y < TH; ++y) s_side[w][o + y] = d[y];
                }
                __syncthreads();
                int nU = hU, nD = hD, nS = hs;
                if (inU) nU = s_bot[w - BX][lane];
                if (inD) nD = s_top[w + BX][lane];
                if (lane < TH && inL) nS = s_side[w - 1][TH + lane];
                if (lane >= TH && lane < 2 * TH && inR) nS = s_side[w + 1][lane - TH];
                const bool hch = have && __any(nU < hU || nD < hD || nS < hs);
                hU = nU;
                hD = nD;
                hs = nS;
                need = hch || !conv;
                // Also the barrier before the next round's LDS writes.
                if (!__syncthreads_or(need)) break;
                if (xr + 1 >= GR_MAX_EXCHANGES) {
                    capped = true;
                    break;
                }
            }
        }
        // A relaxation sweep only lowers distances, so the halo loaded at the
        // start bounds each neighbour's current values from above: a
        // neighbour whose boundary nodes v all have d(u) + 1 >= halo(v) for
        // the changed nodes u next to them cannot improve, and is not queued
        // again. (In a REINIT sweep neighbours may be reset to higher values
        // after the halo was read, so every change counts.) Computed before
        // the stores, while the halos are still in registers.
        bool impL = true, impR = true, impU = true, impD = true;
        if (NB == 4 && NW == 1 && !reinit && have) {
            impU = __any((chg & 1u) && d[0] + 1 < hU);
            impD = __any(((chg >> (TH - 1)) & 1u) && d[TH - 1] + 1 < hD);
            // Every lane compares its rows with the side halo's; lanes 0 and
            // 63 are the boundary columns.
            // (Bitwise: short-circuit && compiled to a branch per row.)
            uint32_t cl = 0, cr = 0;
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                const uint32_t c = (chg >> y) & 1u;
                cl |= c & (uint32_t)(d[y] + 1 < __builtin_amdgcn_readlane(hs, y));
                cr |= c & (uint32_t)(d[y] + 1 < __builtin_amdgcn_readlane(hs, TH + y));
            }
            impL = __any(lane == 0 && cl);
            impR = __any(lane == TW - 1 && cr);
        }
        int act = 0, fin = 0, nearv = 0, open = 0;
        if (have) {
            const int X = tx * TW + lane;
            const bool colin = X > 0 && X < D.W - 1;
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                if (reinit || ((chg >> y) & 1)) T.d[y * TW + lane] = d[y];
                act |= (m.get(y) & MB_EPOS) && d[y] < vinit;
                fin |= d[y] < DINF;
                nearv |= d[y] < near_thr;
                const int Y = ty * TH + y;
                // an interior node not known to reach a sink (at or past the cap, or not at 0 in a flood)
                open |= (int)colin & (int)(Y > 0) & (int)(Y < D.H - 1) & (int)(d[y] >= (flood ? 1 : capv));
            }
            // An exact or reachability round writes the labels of every tile
            // it visits; the last visit of the last round leaves the final ones.
            if (capv >= DINF || flood) write_labels(D, tx, ty, lane, d);
        }
        // Outer boundary changes re-queue the neighbouring blocks.
        bool chL = __any(lane == 0 && chg != 0) && !inL && have && ti.nbr[0] >= 0;
        bool chR = __any(lane == TW - 1 && chg != 0) && !inR && have && ti.nbr[1] >= 0;
        bool chU = __any((chg & 1u) != 0) && !inU && have && ti.nbr[2] >= 0;
        bool chD = __any((chg >> (TH - 1)) & 1u) && !inD && have && ti.nbr[3] >= 0;
        chL = chL && impL;
        chR = chR && impR;
        chU = chU && impU;
        chD = chD && impD;
        // 8-neighbourhood (one-tile blocks): a changed corner node is a halo
        // node of the diagonal tile as well.
        bool chUL = false, chUR = false, chDL = false, chDR = false;
        if (NB == 8 && have) {
            chUL = __any(lane == 0 && (chg & 1u)) && tile_ul(ti) >= 0;
            chUR = __any(lane == TW - 1 && (chg & 1u)) && tile_ur(ti) >= 0;
            chDL = __any(lane == 0 && ((chg >> (TH - 1)) & 1u)) && tile_dl(ti) >= 0;
            chDR = __any(lane == TW - 1 && ((chg >> (TH - 1)) & 1u)) && tile_dr(ti) >= 0;
        }
        const bool any_act = __any(act);
        if (lane == 0 && have) {
            // Block anchors from the compile-time block shape (no divisions).
            auto anchor = [&](int x, int y) { return D.tile0 + (y - y % BY) * D.tiles_x + (x - x % BX); };
            if (chL) mark_dirty(p, anchor(tx - 1, ty));
            if (chR) mark_dirty(p, anchor(tx + 1, ty));
            if (chU) mark_dirty(p, anchor(tx, ty - 1));
            if (chD) mark_dirty(p, anchor(tx, ty + 1));
            if (chUL) mark_dirty(p, tile_ul(ti));
            if (chUR) mark_dirty(p, tile_ur(ti));
            if (chDL) mark_dirty(p, tile_dl(ti));
            if (chDR) mark_dirty(p, tile_dr(ti));
            set_active(p, t, any_act);
        }
        const bool any_fin = NW == 1 ? __any(fin) : __syncthreads_or(fin);
        const bool any_near = p.gr_cap <= 0 || (NW == 1 ? __any(nearv) : __syncthreads_or(nearv));
        const bool any_open = NW == 1 ? __any(open) : __syncthreads_or(open);
        if (leader) {
            if (capped || self_dirty) mark_dirty(p, t0);
            p.live[t0] = any_fin ? ((capv < DINF && !any_open) ? 3 : (any_near ? 2 : 1)) : 0;
        }
    }
    if (p.profile && lane == 0) atomic_add64(&p.ctl->gr_iters, n_rounds);
}

// 8-neighbourhood (NB == 8): every node also holds the diagonal capacities
// {up-left, up-right, down-left, down-right} (KParams::caps8, mask bits in
// KParams::mask8); tiles run in four colours, so no two running tiles touch,
// corners included; instances are never lazy (the build writes every
// capacity, and flow into another tile adds its reverse arcs there).
template <bool ONE, int NB = 4>
__global__ __launch_bounds__(64) void k_prw(KParams p, int parity, GCFG_RO_ARGS) {
    const int qlen = p.ctl->q[parity].qlen_pr;
    if ((int)blockIdx.x >= qlen) return;
    const int lane = threadIdx.x;
    int64_t n_iters = 0;
    // Each workgroup takes a contiguous run of the queue: consecutive items
    // are usually neighbouring tiles of one instance, so their queue words,
    // tile records and descriptor hit the scalar cache, and their halo rows
    // are each other's tile rows.
    const int per = (qlen + (int)gridDim.x - 1) / (int)gridDim.x;
    const int q_end = min(qlen, ((int)blockIdx.x + 1) * per);
    for (int q = (int)blockIdx.x * per; q < q_end; ++q) {
        const int item = rq[q];
        const int t = item & (PR_FRESH - 1);
        const bool fresh = (item & PR_FRESH) != 0;
        const TileRef T = tile_ref(p, t);
        const TileInfo ti = rtiles[t];
        // Descriptor fields and the cap read once, before any store.
        const int dthr = rdesc[ti.inst].n_interior;
        const int tile0 = rdesc[ti.inst].tile0, tiles_x = rdesc[ti.inst].tiles_x, tiles_y = rdesc[ti.inst].tiles_y;
        const int capv = rstate[ti.inst].cap;  // nodes at the cap are frozen until the next relabel
        const bool lz = rdesc[ti.inst].lazy != 0;
        int hU, hD, hs;
        load_halo<HALO_DIST, NB>(p, ti, t, lane, hU, hD, hs);
        int4 c[TH], c8[TH];
        int e[TH], d[TH];
        // Flow pushed out of the tile sideways, one row per lane: lanes 0..15
        // into the left tile, lanes 16..31 into the right tile. With diagonals
        // accSa holds what reached those nodes from the row below them (up-left
        // / up-right arcs) and accSb from the row above (down-left /
        // down-right), and lanes 32..35 the corner nodes' flow (the side-halo
        // lane order).
        int accS = 0, accSa = 0, accSb = 0;
        // Capacities are in the arena (written by the build or by an earlier
        // push pass), or for a fresh tile of a lazy build still only in the
        // inputs; this pass writes them to the arena either way. Excess and
        // distances are loaded first: issued after the branch, the compiler
        // kept both branches' loads in flight at once and the kernel's
        // registers rose past two wavefronts per SIMD.
#pragma unroll
        for (int y = 0; y < TH; ++y) {
            e[y] = T.e[y * TW + lane];
            d[y] = T.d[y * TW + lane];
        }
        if (fresh) {
            const InstDesc D = rdesc[ti.inst];
#pragma unroll
            for (int y = 0; y < TH; ++y) c[y] = input_caps(D, ti.tx * TW + lane, ti.ty * TH + y);
        } else {
#pragma unroll
            for (int y = 0; y < TH; ++y) c[y] = T.c[y * TW + lane];
        }
        if (NB == 8) {
#pragma unroll
            for (int y = 0; y < TH; ++y) c8[y] = p.caps8[(size_t)t * TN + y * TW + lane];
        }
        // Lazy instance: reverse-arc capacity that neighbours' pushes left for
        // the boundary nodes (they do not run in this colour pass): add it,
        // clear the slots.
        if (lz) {
            int32_t *f = p.inflow + (size_t)t * INFLOW;
            const int fs = lane < 2 * TH ? f[lane] : 0;  // lanes 0..15: column 0, 16..31: column TW-1
            const int ft = f[2 * TH + lane], fb = f[2 * TH + TW + lane];
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                const int l = __builtin_amdgcn_readlane(fs, y), r = __builtin_amdgcn_readlane(fs, TH + y);
                c[y].x += lane == 0 ? l : 0;
                c[y].y += lane == TW - 1 ? r : 0;
            }
            c[0].z += ft;
            c[TH - 1].w += fb;
            if (lane < 2 * TH && fs) f[lane] = 0;
            if (ft) f[2 * TH + lane] = 0;
            if (fb) f[2 * TH + TW + lane] = 0;
        }
        // Flow into the rows above and below the tile: straight (u / d arcs),
        // and with diagonals Ua / Da from lane x + 1 (up-left / down-left
        // arcs) and Ub / Db from lane x - 1 (up-right / down-right arcs).
        int accU = 0, accD = 0, accUa = 0, accUb = 0, accDa = 0, accDb = 0;
        // Row y's heights seen one lane over: lane x <- lane x - 1 (lane 0: the
        // left halo column's row y) and lane x <- lane x + 1 (lane 63: the
        // right column's). With diagonals they are also the up-left / up-right
        // heights of row y + 1 and the down-left / down-right ones of row y - 1,
        // so the callers below compute each row's once and pass it on; the
        // halo rows' (with the corner nodes at lanes 0 and 63) are fixed per tile.
        auto shl_row = [&](const int y) {
            return __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(hs, y), d[y], 0x138, 0xF, 0xF, false);
        };
        auto shr_row = [&](const int y) {
            return __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(hs, TH + y), d[y], 0x130, 0xF, 0xF, false);
        };
        int hUL = 0, hUR = 0, hDL = 0, hDR = 0;
        if (NB == 8) {
            hUL = __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(hs, HS_UL), hU, 0x138, 0xF, 0xF, false);
            hUR = __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(hs, HS_UR), hU, 0x130, 0xF, 0xF, false);
            hDL = __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(hs, HS_DL), hD, 0x138, 0xF, 0xF, false);
            hDR = __builtin_amdgcn_update_dpp(__builtin_amdgcn_readlane(hs, HS_DR), hD, 0x130, 0xF, 0xF, false);
        }
        // Push along admissible arcs (d(u) == d(v) + 1) in the order l, r, u,
        // d (then up-left, up-right, down-left, down-right). Heights are those
        // at the start of the iteration; flow received from a lower row can be
        // pushed on in the same sweep.
        // 8-neighbourhood: (hl, hr) are row y's shifted heights, (hul, hur) row
        // y - 1's and (hdl, hdr) row y + 1's; the 4-neighbourhood computes
        // (hl, hr) itself, only for rows it does not skip.
        auto push_row = [&](const int y, int hl, int hr, int hul, int hur, int hdl, int hdr) {
            // A row with no active node pushes nothing (and receives nothing
            // from within itself): skipped whole, a uniform branch.
            if (!__any(e[y] > 0 && d[y] < capv)) return;
            if (NB == 4) {
                hl = shl_row(y);
                hr = shr_row(y);
            }
            const int hu = y ? d[y - 1] : hU;
            const int hd = y < TH - 1 ? d[y + 1] : hD;
            int fl = 0, fr = 0, fu = 0, fd = 0, ful = 0, fur = 0, fdl = 0, fdr = 0;
            int ek = e[y];
            int4 ck = c[y];
            int4 c8k = NB == 8 ? c8[y] : make_int4(0, 0, 0, 0);
            const bool active = ek > 0 && d[y] < capv;
            if (NB == 4) {
                // As selects (nested ifs compiled to exec-mask branches):
                // residual capacities are never negative, so an admissible
                // arc of an active node carries min(excess, capacity), which
                // is 0 once the excess is gone or the arc is saturated.
                const int tgt = d[y] - 1;
                fl = (active && hl == tgt) ? min(ek, ck.x) : 0;
                ek -= fl;
                ck.x -= fl;
                fr = (active && hr == tgt) ? min(ek, ck.y) : 0;
                ek -= fr;
                ck.y -= fr;
                fu = (active && hu == tgt) ? min(ek, ck.z) : 0;
                ek -= fu;
                ck.z -= fu;
                fd = (active && hd == tgt) ? min(ek, ck.w) : 0;
                ek -= fd;
                ck.w -= fd;
            } else if (active) {
                const int tgt = d[y] - 1;
                if (ck.x > 0 && hl == tgt) {
                    fl = min(ek, ck.x);
                    ek -= fl;
                    ck.x -= fl;
                }
                if (ek > 0 && ck.y > 0 && hr == tgt) {
                    fr = min(ek, ck.y);
                    ek -= fr;
                    ck.y -= fr;
                }
                if (ek > 0 && ck.z > 0 && hu == tgt) {
                    fu = min(ek, ck.z);
                    ek -= fu;
                    ck.z -= fu;
                }
                if (ek > 0 && ck.w > 0 && hd == tgt) {
                    fd = min(ek, ck.w);
                    ek -= fd;
                    ck.w -= fd;
                }
                {
                    if (ek > 0 && c8k.x > 0 && hul == tgt) {
                        ful = min(ek, c8k.x);
                        ek -= ful;
                        c8k.x -= ful;
                    }
                    if (ek > 0 && c8k.y > 0 && hur == tgt) {
                        fur = min(ek, c8k.y);
                        ek -= fur;
                        c8k.y -= fur;
                    }
                    if (ek > 0 && c8k.z > 0 && hdl == tgt) {
                        fdl = min(ek, c8k.z);
                        ek -= fdl;
                        c8k.z -= fdl;
                    }
                    if (ek > 0 && c8k.w > 0 && hdr == tgt) {
                        fdr = min(ek, c8k.w);
                        ek -= fdr;
                        c8k.w -= fdr;
                    }
                }
            }
            const int inL = shr1(fr), inR = shl1(fl);
            ck.x += inL;
            ck.y += inR;
            c[y] = ck;
            e[y] = ek + inL + inR;
            const int sfl = __builtin_amdgcn_readlane(fl, 0), sfr = __builtin_amdgcn_readlane(fr, TW - 1);
            accS += lane == y ? sfl : (lane == TH + y ? sfr : 0);
            if (NB == 8) {
                c8[y] = c8k;
                // Diagonal flow out of lanes 0 / 63 enters the side columns one
                // row up or down (a corner tile past the first or last row).
                const int sful = __builtin_amdgcn_readlane(ful, 0), sfur = __builtin_amdgcn_readlane(fur, TW - 1);
                const int sfdl = __builtin_amdgcn_readlane(fdl, 0), sfdr = __builtin_amdgcn_readlane(fdr, TW - 1);
                if (y > 0) accSa += lane == y - 1 ? sful : (lane == TH + y - 1 ? sfur : 0);
                else accSa += lane == HS_UL ? sful : (lane == HS_UR ? sfur : 0);
                if (y < TH - 1) accSb += lane == y + 1 ? sfdl : (lane == TH + y + 1 ? sfdr : 0);
                else accSb += lane == HS_DL ? sfdl : (lane == HS_DR ? sfdr : 0);
                // Inside the tile: lane x - 1 / x + 1 of the row above and below
                // gains the flow and the reverse arc (down-right / down-left
                // for up-left / up-right pushes, up-right / up-left for
                // down-left / down-right ones).
                const int iul = shl1(ful), iur = shr1(fur), idl = shl1(fdl), idr = shr1(fdr);
                if (y > 0) {
                    e[y - 1] += iul + iur;
                    c8[y - 1].w += iul;
                    c8[y - 1].z += iur;
                } else {
                    accUa += iul;
                    accUb += iur;
                }
                if (y < TH - 1) {
                    e[y + 1] += idl + idr;
                    c8[y + 1].y += idl;
                    c8[y + 1].x += idr;
                } else {
                    accDa += idl;
                    accDb += idr;
                }
            }
            if (y > 0) {
                e[y - 1] += fu;
                c[y - 1].w += fu;
            } else {
                accU += fu;
            }
            if (y < TH - 1) {
                e[y + 1] += fd;
                c[y + 1].z += fd;
            } else {
                accD += fd;
            }
        };
        // Relabel a node left with excess: d = 1 + min{d(v) : r(u,v) > 0}.
        // Rows are relabelled in order, each with the new height of the row
        // above and the old one of the row below. Returns whether the node
        // is still active (below the cap).
        // (8-neighbourhood: hul, hur from the new heights of row y - 1, hdl,
        // hdr from the old ones of row y + 1.)
        auto relabel_row = [&](const int y, int hl, int hr, int hul, int hur, int hdl, int hdr) -> int {
            if (!__any(e[y] > 0 && d[y] < capv)) return 0;
            if (NB == 4) {
                hl = shl_row(y);
                hr = shr_row(y);
            }
            const int hu = y ? d[y - 1] : hU;
            const int hd = y < TH - 1 ? d[y + 1] : hD;
            // Every lane computes; only active nodes take the new label (a
            // per-lane early return compiled to an exec-mask branch).
            const bool active = e[y] > 0 && d[y] < capv;
            int mm = DINF;
            if (c[y].x > 0) mm = min(mm, hl);
            if (c[y].y > 0) mm = min(mm, hr);
            if (c[y].z > 0) mm = min(mm, hu);
            if (c[y].w > 0) mm = min(mm, hd);
            if (NB == 8) {
                if (c8[y].x > 0) mm = min(mm, hul);
                if (c8[y].y > 0) mm = min(mm, hur);
                if (c8[y].z > 0) mm = min(mm, hdl);
                if (c8[y].w > 0) mm = min(mm, hdr);
            }
            int nd = mm >= DINF ? DINF : mm + 1;
            // A label of n or more proves the node cannot reach the sink; a
            // label past the cap freezes the node at the cap (still a valid
            // labelling: cap < nd <= d(v) + 1).
            if (nd >= dthr) nd = DINF;
            else if (nd > capv) nd = capv;
            const int raised = max(d[y], nd);
            d[y] = active ? raised : d[y];
            asm volatile("" : "+v"(d[y]));  // kept a select (as an if it became a branch)
            return active && d[y] < capv;
        };
        int mine = 0;
        int64_t sx = 0;  // sum of max(e, 0) over the tile (the flow identity's stuck excess)
        auto store_row = [&](const int y) {
            T.c[y * TW + lane] = c[y];
            T.e[y * TW + lane] = e[y];
            T.d[y * TW + lane] = d[y];
            T.m[y * TW + lane] = (uint8_t)node_mask(c[y], e[y]);
            if (NB == 8) {
                const size_t g = (size_t)t * TN + y * TW + lane;
                p.caps8[g] = c8[y];
                p.mask8[g] = (uint8_t)node_mask(c8[y], 0);
            }
            mine |= e[y] > 0 && d[y] < capv;
            sx += max(e[y], 0);
        };
        int it = 0;
        // Shifted heights carried along the rows (8-neighbourhood only): pl/pr
        // row y - 1's, cl/cr row y's, nl/nr row y + 1's, all as of the start
        // of the push phase; npl/npr the relabelled row above's.
        int pl = hUL, pr = hUR, cl = 0, cr = 0, nl = 0, nr = 0, npl = hUL, npr = hUR;
        auto next_row = [&](const int y) {  // nl/nr <- row y + 1 (the halo row below past the last)
            if (NB != 8) return;
            nl = y < TH - 1 ? shl_row(y + 1) : hDL;
            nr = y < TH - 1 ? shr_row(y + 1) : hDR;
        };
        if (ONE) {
            // One iteration: row y - 1 is final once row y has pushed (only
            // rows y - 2 .. y push into it), so it is relabelled and stored
            // right then, with exactly the values of the general order, and
            // its registers are free for the rows still to come.
            it = 1;
            if (NB == 8) {
                cl = shl_row(0);
                cr = shr_row(0);
            }
#pragma unroll
            for (int y = 0; y < TH; ++y) {
                next_row(y);
                push_row(y, cl, cr, pl, pr, nl, nr);
                if (y > 0) {
                    relabel_row(y - 1, pl, pr, npl, npr, cl, cr);
                    if (NB == 8) {
                        npl = shl_row(y - 1);
                        npr = shr_row(y - 1);
                    }
                    store_row(y - 1);
                }
                pl = cl;
                pr = cr;
                cl = nl;
                cr = nr;
            }
            relabel_row(TH - 1, pl, pr, npl, npr, hDL, hDR);
            store_row(TH - 1);
        } else {
            // Exact rounds come after the capped ones when flow still has far to
            // go: their passes push for twice as many iterations (a 4096^2
            // image with heavy-tailed t-links 7.1 -> 6.6 ms; instances that
            // finish within the capped rounds never run one).
            const int iters = capv >= DINF ? min(2 * p.inner_iters, 64) : p.inner_iters;
            // One iteration: pushes with the heights of its start, then
            // relabels. 4-neighbourhood: every other iteration pushes from the
            // last row up, so flow pushed into the row above is pushed on in
            // the same iteration, as flow into the row below is on the way
            // down (either order is a legal sequence of pushes). Returns
            // whether a node is still active.
            auto iteration = [&](auto upward) -> bool {
                if (NB == 8) {
                    pl = hUL;
                    pr = hUR;
                    cl = shl_row(0);
                    cr = shr_row(0);
                }
                if (NB == 4 && decltype(upward)::value) {
#pragma unroll
                    for (int y = TH - 1; y >= 0; --y) push_row(y, 0, 0, 0, 0, 0, 0);
                } else {
#pragma unroll
                    for (int y = 0; y < TH; ++y) {
                        next_row(y);
                        push_row(y, cl, cr, pl, pr, nl, nr);
                        pl = cl;
                        pr = cr;
                        cl = nl;
                        cr = nr;
                    }
                }
                int act = 0;
                if (NB == 8) {
                    npl = hUL;
                    npr = hUR;
                    cl = shl_row(0);
                    cr = shr_row(0);
                }
#pragma unroll
                for (int y = 0; y < TH; ++y) {
                    next_row(y);
                    act |= relabel_row(y, cl, cr, npl, npr, nl, nr);
                    if (NB == 8) {
                        npl = shl_row(y);
                        npr = shr_row(y);
                    }
                    cl = nl;
                    cr = nr;
                }
                return __any(act);
            };
            while (it < iters) {
                ++it;
                if (!iteration(std::false_type{})) break;
                if (NB == 8 || it >= iters) continue;
                ++it;
                if (!iteration(std::true_type{})) break;
            }
#pragma unroll
            for (int y = 0; y < TH; ++y) store_row(y);
        }
        // Flow pushed into halo nodes (their tiles do not run in this colour
        // pass; int32 atomics because two same-colour tiles can feed one
        // corner node). Lanes 0..31 feed the side tiles (one row each), lanes
        // 32..35 the corner tiles (8-neighbourhood), every lane may feed the
        // tiles above and below. All excess atomics are issued before their
        // results are used for the mask updates, so their latencies overlap.
        int sidx;
        const int side = side_halo<NB>(ti, lane, sidx);
        const int totS = accS + accSa + accSb, totU = accU + accUa + accUb, totD = accD + accDa + accDb;
        const bool pushS = totS != 0 && side >= 0;
        const size_t gS = (size_t)(pushS ? side : 0) * TN + (size_t)sidx;
        // Receiving nodes left of the tile: the left column and the left corners.
        const bool leftS = lane < TH || lane == HS_UL || lane == HS_DL;
        // The side node's arc back to us: the right arc of the left tile's
        // column TW-1 (slot TH + row), the left arc of the right tile's column 0 (slot row).
        const int slotS = lane < TH ? TH + lane : lane - TH;
        const int arcS = lane < TH ? 2 : 1;
        const bool pushU = totU != 0 && ti.nbr[2] >= 0, pushD = totD != 0 && ti.nbr[3] >= 0;
        const size_t gU = (size_t)(pushU ? ti.nbr[2] : 0) * TN + (size_t)(TH - 1) * TW + lane;
        const size_t gD = (size_t)(pushD ? ti.nbr[3] : 0) * TN + lane;
        int oldS = 0, oldU = 0, oldD = 0;
        // Reverse arcs: in a lazy instance to the receiving tile's inflow
        // slots (its capacities may still live in the inputs; down arcs of the
        // upper tile's last row, up arcs of the lower tile's first row),
        // otherwise to its capacities directly. Diagonal reverse arcs (never
        // lazy) go to the diagonal capacities {ul, ur, dl, dr}: flow from the
        // row below a node opens its down-left or down-right arc, flow from
        // the row above its up-left or up-right arc.
        int *revS = lz ? &p.inflow[(size_t)(pushS ? side : 0) * INFLOW + slotS]
                       : reinterpret_cast<int *>(&p.caps[gS]) + (lane < TH ? 1 : 0);
        int *revU = lz ? &p.inflow[(size_t)(pushU ? ti.nbr[2] : 0) * INFLOW + 2 * TH + TW + lane]
                       : reinterpret_cast<int *>(&p.caps[gU]) + 3;
        int *revD = lz ? &p.inflow[(size_t)(pushD ? ti.nbr[3] : 0) * INFLOW + 2 * TH + lane]
                       : reinterpret_cast<int *>(&p.caps[gD]) + 2;
        if (pushS) {
            oldS = atomicAdd(&p.exc[gS], totS);
            if (accS) atomicAdd(revS, accS);
            if (NB == 8) {
                int *r8 = reinterpret_cast<int *>(&p.caps8[gS]);
                if (accSa) atomicAdd(r8 + (leftS ? 3 : 2), accSa);
                if (accSb) atomicAdd(r8 + (leftS ? 1 : 0), accSb);
            }
        }
        if (pushU) {
            oldU = atomicAdd(&p.exc[gU], totU);
            if (accU) atomicAdd(revU, accU);
            if (NB == 8) {
                int *r8 = reinterpret_cast<int *>(&p.caps8[gU]);
                if (accUa) atomicAdd(r8 + 3, accUa);
                if (accUb) atomicAdd(r8 + 2, accUb);
            }
        }
        if (pushD) {
            oldD = atomicAdd(&p.exc[gD], totD);
            if (accD) atomicAdd(revD, accD);
            if (NB == 8) {
                int *r8 = reinterpret_cast<int *>(&p.caps8[gD]);
                if (accDa) atomicAdd(r8 + 1, accDa);
                if (accDb) atomicAdd(r8 + 0, accDb);
            }
        }
        if (pushS) mask_after_push_in(p, gS, accS ? arcS : 0, oldS + totS);
        if (pushU) mask_after_push_in(p, gU, accU ? 8 : 0, oldU + totU);
        if (pushD) mask_after_push_in(p, gD, accD ? 4 : 0, oldD + totD);
        if (NB == 8) {
            const int bS = (accSa ? (leftS ? 8 : 4) : 0) | (accSb ? (leftS ? 2 : 1) : 0);
            const int bU = (accUa ? 8 : 0) | (accUb ? 4 : 0), bD = (accDa ? 2 : 0) | (accDb ? 1 : 0);
            if (pushS && bS) mask8_after_push_in(p, gS, bS);
            if (pushU && bU) mask8_after_push_in(p, gU, bU);
            if (pushD && bD) mask8_after_push_in(p, gD, bD);
        }
        // Stuck-excess sums of the tiles pushed into (they are not running).
        if (pushS) atomic_add64(&p.tsum[2 * (size_t)side + 1], (int64_t)max(oldS + totS, 0) - max(oldS, 0));
        if (pushU) atomic_add64(&p.tsum[2 * (size_t)ti.nbr[2] + 1], (int64_t)max(oldU + totU, 0) - max(oldU, 0));
        if (pushD) atomic_add64(&p.tsum[2 * (size_t)ti.nbr[3] + 1], (int64_t)max(oldD + totD, 0) - max(oldD, 0));
        sx = wave_sum64_dpp(sx);
        const int toL = pushS && lane < TH, toR = pushS && lane >= TH && lane < 2 * TH;
        const bool anyL = __any(toL), anyR = __any(toR);
        const bool anyU = __any(pushU), anyD = __any(pushD);
        // 8-neighbourhood: the corner lanes' receiving tiles (side-halo order).
        const uint64_t corners = NB == 8 ? __ballot(pushS && lane >= HS_UL) >> HS_UL : 0;
        const bool any_mine = __any(mine);
        if (lane == 0) {
            if (anyL) set_active(p, ti.nbr[0], 1);
            if (anyR) set_active(p, ti.nbr[1], 1);
            if (anyU) set_active(p, ti.nbr[2], 1);
            if (anyD) set_active(p, ti.nbr[3], 1);
            if (NB == 8) {
                if (corners & 1) set_active(p, tile_ul(ti), 1);
                if (corners & 2) set_active(p, tile_ur(ti), 1);
                if (corners & 4) set_active(p, tile_dl(ti), 1);
                if (corners & 8) set_active(p, tile_dr(ti), 1);
            }
            set_active(p, t, any_mine);
            p.tsum[2 * (size_t)t + 1] = sx;
            if (fresh) p.mat[t] = 1;
        }
        if (p.gr_cap > 0) {
            // Blocks within cap distance of this tile (and of the tiles it
            // pushed into) need their capped distances recomputed. A path of
            // length K moves at most K columns and K rows, diagonals included.
            const int rx = (p.gr_cap_late + TW - 1) / TW + ((anyL || anyR || corners) ? 1 : 0);
            const int ry = (p.gr_cap_late + TH - 1) / TH + ((anyU || anyD || corners) ? 1 : 0);
            const int span = 2 * rx + 1, cnt = span * (2 * ry + 1);
            for (int k = lane; k < cnt; k += TW) {
                const int x = ti.tx + k % span - rx, y = ti.ty + k / span - ry;
                if (x >= 0 && y >= 0 && x < tiles_x && y < tiles_y)
                    p.stale[(p.gbx == 1 && p.gby == 1) ? tile0 + y * tiles_x + x
                                                       : tile0 + (y - y % p.gby) * tiles_x + (x - x % p.gbx)] = 1;
            }
        }
        n_iters += it;
    }
    if (p.profile && lane == 0) atomic_add64(&p.ctl->pr_iters, n_iters);
}

// ---------------------------------------------------------------------------
// Wavefront build (GC_SetGraph) fused with the first global-relabel sweep.
__device__ __forceinline__ int find_inst(const int32_t *prefix, int n, int b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Tile table and cleared flags of the listed instances (one thread per tile).
// Tile record, map entry and flags of build-list position b.
__device__ void clear_tile(const KParams &p, const int32_t *slots, const int32_t *prefix, int n, int b) {
    const int j = find_inst(prefix, n, b);
    const int s = slots[j];
    const InstDesc D = p.desc[s];
    const int lt = b - prefix[j];
    const int t = D.tile0 + lt;
    const int tx = lt % D.tiles_x, ty = lt / D.tiles_x;
    TileInfo ti;
    ti.inst = s;
    ti.tx = tx;
    ti.ty = ty;
    ti.nbr[0] = tx > 0 ? t - 1 : -1;
    ti.nbr[1] = tx < D.tiles_x - 1 ? t + 1 : -1;
    ti.nbr[2] = ty > 0 ? t - D.tiles_x : -1;
    ti.nbr[3] = ty < D.tiles_y - 1 ? t + D.tiles_x : -1;
    ti.pad = 0;
    p.tiles[t] = ti;
    p.act_flag[t] = 0;
    p.dirty[t] = 0;
    p.live[t] = 0;  // raised by the build sweep of the block's tiles
    p.stale[t] = 0;
    p.bmap[b] = t;
}


// Initial BFS distance of an input pixel: 0 if it is an interior node with
// pos < neg (remaining sink capacity), otherwise unreachable so far.
__device__ __forceinline__ int input_d0(const InstDesc &D, int X, int Y) {
    const bool in = X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1;
    const size_t i = in ? (size_t)Y * D.W + X : 0;  // clamped: the load is always in bounds
    const int pp = ldg_in(D.pos, i), qq = ldg_in(D.neg, i);
    return (in && pp < qq) ? 0 : DINF;
}

// The build map, tile records and descriptors come in as restrict read-only
// arguments, like the solve kernels' queue arguments: their loads are uniform
// and go through the scalar cache, so a tile's prologue is one scalar round
// trip per level instead of vector loads and waits.
template <int NB = 4>
__global__ __launch_bounds__(64) void k_buildw(KParams p, const int32_t *__restrict__ rbmap,
                                               const TileInfo *__restrict__ rtiles,
                                               const InstDesc *__restrict__ rdesc, int total) {
    const int lane = threadIdx.x;
    // Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.
    // Giving every XCD a contiguous eighth of the tiles keeps horizontally
    // and vertically adjacent tiles, which share input cache lines (rows of
    // the caller's row-major planes are not 256-byte aligned), in one L2.
    const bool by_xcd = gridDim.x % NUM_XCD == 0;
    const int per = (total + NUM_XCD - 1) / NUM_XCD;
    const int jstep = by_xcd ? gridDim.x / NUM_XCD : gridDim.x;
    const int jend = by_xcd ? per : total;
    const int b0 = by_xcd ? (int)(blockIdx.x % NUM_XCD) * per : 0;
    for (int jj = by_xcd ? blockIdx.x / NUM_XCD : blockIdx.x; jj < jend; jj += jstep) {
        const int b = b0 + jj;
        if (b >= total) break;
        // Tile, instance and tile coordinates from the map and the tile
        // records k_arm wrote just before (two dependent loads).
        const int t = rbmap[b];
        const TileInfo ti = rtiles[t];
        const int s = ti.inst;
        const InstDesc D = rdesc[s];
        const int tx = ti.tx, ty = ti.ty;
        const int X = tx * TW + lane, Y0 = ty * TH;
        const TileRef T = tile_ref(p, t);
        int d[TH];
        Masks m, m8;
        m.clear();
        m8.clear();
        const int capv = p.gr_cap > 0 ? p.gr_cap : DINF;  // the first relabel sweep, capped like k_arm says
        int64_t sp = 0, sx = 0;
        int bad = 0;
        // Halo distances from the inputs (issued first, consumed after the
        // rows; loaded unconditionally, from clamped in-bounds addresses, so
        // that no branch waits for them).
        const int hUv = input_d0(D, X, Y0 - 1), hDv = input_d0(D, X, Y0 + TH);
        const int hU = ty > 0 ? hUv : DINF;
        const int hD = ty < D.tiles_y - 1 ? hDv : DINF;
        const bool hleft = lane < TH && tx > 0, hright = lane >= TH && lane < 2 * TH && tx < D.tiles_x - 1;
        // 8-neighbourhood: lanes 32..35 the corner nodes (up-left, up-right,
        // down-left, down-right), where the diagonal tile exists.
        const bool cl = lane == HS_UL || lane == HS_DL, cu = lane == HS_UL || lane == HS_UR;
        const bool corner = NB == 8 && lane >= HS_UL && lane <= HS_DR && (cl ? tx > 0 : tx < D.tiles_x - 1) &&
                            (cu ? ty > 0 : ty < D.tiles_y - 1);
        const int hsx = lane < 2 * TH ? (lane < TH ? tx * TW - 1 : tx * TW + TW) : (cl ? tx * TW - 1 : tx * TW + TW);
        const int hsy = lane < 2 * TH ? Y0 + (lane & (TH - 1)) : (cu ? Y0 - 1 : Y0 + TH);
        const int hsv = input_d0(D, hsx, hsy);
        const int hs = (hleft || hright || corner) ? hsv : DINF;
        // Two halves of 8 rows: all input loads of a half, then compute and
        // store (clamped in-bounds loads; values outside the interior masked).
#pragma unroll
        for (int h = 0; h < TH; h += TH / 2) {
            int vp[TH / 2], vn[TH / 2];
            int4 vc[TH / 2], vc8[TH / 2];
            bool vin[TH / 2];
#pragma unroll
            for (int k = 0; k < TH / 2; ++k) {
                const int Y = Y0 + h + k;
                const bool in = (X > 0) & (Y > 0) & (X < D.W - 1) & (Y < D.H - 1);
                const size_t i = in ? (size_t)((uint32_t)Y * (uint32_t)D.W + (uint32_t)X) : 0;
                vin[k] = in;
                vp[k] = ldg_in(D.pos, i);
                vn[k] = ldg_in(D.neg, i);
                if (D.up)
                    vc[k] = make_int4(ldg_in(D.left, i), ldg_in(D.right, i), ldg_in(D.up, i), ldg_in(D.down, i));
                else
                    vc[k] = make_int4(D.penalty, D.penalty, D.penalty, D.penalty);
                if (NB == 8) {
                    if (D.up)
                        vc8[k] = make_int4(ldg_in(D.upleft, i), ldg_in(D.upright, i), ldg_in(D.downleft, i),
                                           ldg_in(D.downright, i));
                    else
                        vc8[k] = make_int4(D.penalty, D.penalty, D.penalty, D.penalty);
                }
            }
#pragma unroll
            for (int k = 0; k < TH / 2; ++k) {
                const int y = h + k, Y = Y0 + y;
                int4 c = make_int4(0, 0, 0, 0), c8 = make_int4(0, 0, 0, 0);
                int e = 0;
                if (vin[k]) {
                    c = vc[k];
                    if (X == 1) c.x = 0;
                    if (X == D.W - 2) c.y = 0;
                    if (Y == 1) c.z = 0;
                    if (Y == D.H - 2) c.w = 0;
                    if (NB == 8) {
                        // Diagonal arcs towards the border are zero like the
                        // side arcs (GraphCutKernels.cu:133-151,190-213).
                        c8 = vc8[k];
                        if (X == 1) c8.x = c8.z = 0;
                        if (X == D.W - 2) c8.y = c8.w = 0;
                        if (Y == 1) c8.x = c8.y = 0;
                        if (Y == D.H - 2) c8.z = c8.w = 0;
                        bad |= nlinks_bad(c8, 8);
                    }
                    bad |= tlinks_bad(vp[k], vn[k]) | nlinks_bad(c, NB);
                    e = vp[k] - vn[k];
                    sp += vp[k];
                    sx += max(e, 0);
                }
                // Lazy build: capacities stay in the inputs until a push pass
                // needs them (most tiles never see one). Excess is not read
                // again before a push pass either: a streaming store.
                if (!D.lazy) T.c[y * TW + lane] = c;
                __builtin_nontemporal_store(e, &T.e[y * TW + lane]);
                const int mk = node_mask(c, e);
                m.put(y, mk);
                T.m[y * TW + lane] = (uint8_t)mk;
                if (NB == 8) {
                    // Never lazy: the diagonal capacities are written here.
                    const size_t g = (size_t)t * TN + y * TW + lane;
                    const int mk8 = node_mask(c8, 0);
                    m8.put(y, mk8);
                    p.caps8[g] = c8;
                    p.mask8[g] = (uint8_t)mk8;
                }
                d[y] = e < 0 ? 0 : capv;
            }
        }
        {
            int32_t *f = p.inflow + (size_t)t * INFLOW;
            if (lane < 2 * TH) f[lane] = 0;
            f[2 * TH + lane] = 0;
            f[2 * TH + TW + lane] = 0;
        }
        uint32_t chg = 0;
        int rounds = 0;
        // Sources: the sinks and the halo sinks read from the inputs, all at 0.
        const bool conv = tile_bfs<NB>(d, m, m8, hU, hD, hs, lane, capv, chg, rounds, p.gr_rounds, true);
        int act = 0, fin = 0, nearv = 0, open = 0;
        const bool colin = X > 0 && X < D.W - 1;
#pragma unroll
        for (int y = 0; y < TH; ++y) {
            T.d[y * TW + lane] = d[y];
            act |= (m.get(y) & MB_EPOS) && d[y] < capv;
            fin |= d[y] < DINF;
            nearv |= d[y] < capv;
            open |= (int)colin & (int)(Y0 + y > 0) & (int)(Y0 + y < D.H - 1) & (int)(d[y] >= capv);
        }
        // Uncapped, this sweep belongs to an exact relabel round: labels as
        // of this sweep (rewritten by any later sweep of the tile).
        if (capv >= DINF) write_labels(D, tx, ty, lane, d);
        // Liveness as k_grb keeps it; 3 (every interior node below the cap)
        // only for one-tile blocks, whose rank is the tile's own.
        const bool one = p.gbx == 1 && p.gby == 1;
        const int rank = __any(fin) ? ((one && capv < DINF && !__any(open)) ? 3 : (__any(nearv) ? 2 : 1)) : 0;
        const bool chL = __any(lane == 0 && chg != 0);
        const bool chR = __any(lane == TW - 1 && chg != 0);
        const bool chU = __any((chg & 1u) != 0);
        const bool chD = __any((chg >> (TH - 1)) & 1u);
        // 8-neighbourhood: corner nodes are halo nodes of the diagonal tiles.
        const bool chC0 = NB == 8 && __any(lane == 0 && (chg & 1u));
        const bool chC1 = NB == 8 && __any(lane == TW - 1 && (chg & 1u));
        const bool chC2 = NB == 8 && __any(lane == 0 && ((chg >> (TH - 1)) & 1u));
        const bool chC3 = NB == 8 && __any(lane == TW - 1 && ((chg >> (TH - 1)) & 1u));
        const bool any_act = __any(act);
        const bool any_bad = __any(bad);
        sp = wave_sum64_dpp(sp);
        sx = wave_sum64_dpp(sx);
        if (lane == 0) {
            // Dirty marks go to the global-relabel blocks of the affected tiles.
            if (chL && tx > 0) mark_dirty(p, block_anchor(p, D, tx - 1, ty));
            if (chR && tx < D.tiles_x - 1) mark_dirty(p, block_anchor(p, D, tx + 1, ty));
            if (chU && ty > 0) mark_dirty(p, block_anchor(p, D, tx, ty - 1));
            if (chD && ty < D.tiles_y - 1) mark_dirty(p, block_anchor(p, D, tx, ty + 1));
            if (chC0 && tx > 0 && ty > 0) mark_dirty(p, block_anchor(p, D, tx - 1, ty - 1));
            if (chC1 && tx < D.tiles_x - 1 && ty > 0) mark_dirty(p, block_anchor(p, D, tx + 1, ty - 1));
            if (chC2 && tx > 0 && ty < D.tiles_y - 1) mark_dirty(p, block_anchor(p, D, tx - 1, ty + 1));
            if (chC3 && tx < D.tiles_x - 1 && ty < D.tiles_y - 1) mark_dirty(p, block_anchor(p, D, tx + 1, ty + 1));
            if (!conv) mark_dirty(p, block_anchor(p, D, tx, ty));
            set_active(p, t, any_act);
            if (rank) atomicMax(&p.live[block_anchor(p, D, tx, ty)], rank);
            p.tsum[2 * (size_t)t] = sp;
            p.tsum[2 * (size_t)t + 1] = sx;
            p.mat[t] = D.lazy ? 0 : 1;
            if (any_bad) atomicOr(&p.state[s].err, 1);
        }
    }
}

// ---------------------------------------------------------------------------
// Warm-start edit: strokes of (x0, y0, w, h, cls) add delta to pos (fg) or
// neg (bg) of interior pixels, i.e. excess += delta or -= delta.
__global__ __launch_bounds__(NT) void k_edit(KParams p, int slot, const int32_t *strokes, int n, int32_t delta) {
    __shared__ int64_t red[NT / 64];
    const InstDesc D = p.desc[slot];
    for (int s = blockIdx.x; s < n; s += gridDim.x) {
        const int x0 = strokes[5 * s], y0 = strokes[5 * s + 1], w = strokes[5 * s + 2], h = strokes[5 * s + 3];
        const int cls = strokes[5 * s + 4];
        int64_t dp = 0;
        for (int i = threadIdx.x; i < w * h; i += NT) {
            const int X = x0 + i % w, Y = y0 + i / w;
            if (X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1) {
                const int t = D.tile0 + (Y / TH) * D.tiles_x + X / TW;
                const size_t g = (size_t)t * TN + (Y % TH) * TW + (X % TW);
                const int de = cls == GCFG_SEED_FG ? delta : -delta;
                const int old = atomicAdd(&p.exc[g], de);
                if (cls == GCFG_SEED_FG) dp += delta;
                // Stuck-excess sum of the tile (wavefront kernels keep it current).
                const int64_t dx = (int64_t)max(old + de, 0) - max(old, 0);
                if (dx) atomic_add64(&p.tsum[2 * (size_t)t + 1], dx);
            }
        }
        const int64_t tot = block_sum64(dp, red);
        if (threadIdx.x == 0 && tot) atomic_add64(&p.state[slot].sum_pos, tot);
        __syncthreads();
    }
}

// After k_edit: excess-sign bits of the residual masks under the strokes
// (idempotent where strokes overlap).
__global__ __launch_bounds__(NT) void k_edit_masks(KParams p, int slot, const int32_t *strokes, int n) {
    const InstDesc D = p.desc[slot];
    for (int s = blockIdx.x; s < n; s += gridDim.x) {
        const int x0 = strokes[5 * s], y0 = strokes[5 * s + 1], w = strokes[5 * s + 2], h = strokes[5 * s + 3];
        for (int i = threadIdx.x; i < w * h; i += NT) {
            const int X = x0 + i % w, Y = y0 + i / w;
            if (X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1) {
                const int t = D.tile0 + (Y / TH) * D.tiles_x + X / TW;
                const size_t g = (size_t)t * TN + (Y % TH) * TW + (X % TW);
                const int e = p.exc[g];
                p.mask[g] = (uint8_t)((p.mask[g] & 15) | (e > 0 ? MB_EPOS : 0) | (e < 0 ? MB_ENEG : 0));
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Warm start onto new capacities (the next video frame of a session): the
// residual graph of the previous solve is reparametrised to the new inputs
// instead of rebuilt (the reference rebuilds every frame,
// ComplexApplication.cu:127-135,202-207 -> GC_SetGraph).
//  * t-links: excess += (pos1 - neg1) - (pos0 - neg0), and the interior sum
//    of pos moves by pos1 - pos0;
//  * n-link pair (u, v): with old capacities c_uv, c_vu and residuals r_uv,
//    r_vu the net flow u -> v is g = c_uv - r_uv. It is clamped to the new
//    capacities, g' = min(max(g, -c'_vu), c'_uv); then r_uv = c'_uv - g',
//    r_vu = c'_vu + g', and the flow no longer carried, g - g', stays at u
//    (excess_u += g - g', excess_v -= g - g').
// The result is a preflow of the new graph with the same conservation
// identity, so flow = sum pos - sum max(excess, 0) still holds at the end.
// One thread per interior node owns its right and down pairs (each residual
// component has one writer); excess changes are int32 atomics. Old inputs
// come from the descriptor, new ones from the arguments (null up = penalty).
__device__ __forceinline__ int reframe_cap(const int32_t *plane, int32_t pen, size_t i) {
    return plane ? plane[i] : pen;
}

__global__ __launch_bounds__(NT) void k_reframe(KParams p, int slot, const int32_t *pos1, const int32_t *neg1,
                                                const int32_t *up1, const int32_t *down1, const int32_t *left1,
                                                const int32_t *right1, int32_t pen1) {
    __shared__ int64_t red[NT / 64];
    const InstDesc D = p.desc[slot];
    const int64_t n = (int64_t)D.W * D.H;
    int64_t dp = 0;
    int bad = 0;
    auto node = [&](int X, int Y) -> size_t {
        const int t = D.tile0 + (Y / TH) * D.tiles_x + X / TW;
        return (size_t)t * TN + (Y % TH) * TW + (X % TW);
    };
    auto inside = [&](int X, int Y) { return X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1; };
    for (int64_t i = blockIdx.x * (int64_t)NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) {
        const int X = (int)(i % D.W), Y = (int)(i / D.W);
        if (!inside(X, Y)) continue;
        const size_t gu = node(X, Y);
        const int p0 = D.pos[i], q0 = D.neg[i], p1 = pos1[i], q1 = neg1[i];
        bad |= tlinks_bad(p1, q1);
        // The change of a node's excess can leave int32 on its own (t-link
        // differences up to 2^29 each, pair flows up to 2^28); the final
        // excess is in range (see tlinks_bad), so the parts are summed in 64
        // bits and applied with 32-bit atomics, which wrap, in any order.
        int64_t de = ((int64_t)p1 - q1) - ((int64_t)p0 - q0);
        dp += (int64_t)p1 - p0;
        // pairs (u, right neighbour) and (u, lower neighbour)
        for (int k = 0; k < 2; ++k) {
            const int Xv = X + (k == 0), Yv = Y + (k == 1);
            if (!inside(Xv, Yv)) continue;
            const size_t iv = (size_t)Yv * D.W + Xv, gv = node(Xv, Yv);
            // Net flow u -> v from one side of the pair (the pair's residual
            // sum is the sum of its two old capacities).
            const int c_uv0 = reframe_cap(k ? D.down : D.right, D.penalty, i);
            const int c_uv1 = reframe_cap(k ? down1 : right1, pen1, i);
            const int c_vu1 = reframe_cap(k ? up1 : left1, pen1, iv);
            bad |= nlink_bad(c_uv1, 4) | nlink_bad(c_vu1, 4);
            int *ru = reinterpret_cast<int *>(&p.caps[gu]) + (k ? 3 : 1);  // u -> v: down / right
            int *rv = reinterpret_cast<int *>(&p.caps[gv]) + (k ? 2 : 0);  // v -> u: up / left
            const int g = c_uv0 - *ru;
            const int g1 = min(max(g, -c_vu1), c_uv1);
            *ru = c_uv1 - g1;
            *rv = c_vu1 + g1;
            if (g != g1) {
                de += (int64_t)g - g1;
                atomicAdd(reinterpret_cast<unsigned *>(&p.exc[gv]), (unsigned)(uint32_t)(g1 - (int64_t)g));
            }
        }
        if (de) atomicAdd(reinterpret_cast<unsigned *>(&p.exc[gu]), (unsigned)(uint32_t)de);
    }
    const int64_t tot = block_sum64(dp, red);
    const int anybad = __syncthreads_or(bad);
    if (threadIdx.x == 0) {
        if (tot) atomic_add64(&p.state[slot].sum_pos, tot);
        if (anybad) atomicOr(&p.state[slot].err, 1);
    }
}

// After k_reframe: residual masks and the per-tile stuck-excess sums of the
// instance from the new residual graph (one workgroup per tile).
__global__ __launch_bounds__(NT) void k_remask(KParams p, int slot) {
    __shared__ int64_t red[NT / 64];
    const InstDesc D = p.desc[slot];
    for (int lt = blockIdx.x; lt < D.ntiles; lt += gridDim.x) {
        const int t = D.tile0 + lt;
        const int tx = lt % D.tiles_x, ty = lt / D.tiles_x;
        int64_t sx = 0;
        for (int j = threadIdx.x; j < TN; j += NT) {
            const size_t g = (size_t)t * TN + j;
            const int X = tx * TW + j % TW, Y = ty * TH + j / TW;
            const int e = p.exc[g];
            p.mask[g] = (uint8_t)node_mask(p.caps[g], e);
            if (X > 0 && Y > 0 && X < D.W - 1 && Y < D.H - 1) sx += max(e, 0);
        }
        const int64_t tot = block_sum64(sx, red);
        if (threadIdx.x == 0) p.tsum[2 * (size_t)t + 1] = tot;
        __syncthreads();
    }
}

// Write the residual capacities of every tile of an instance to caps (from
// the inputs where a tile has not been pushed since its build) with the
// pending inflow added, then clear the inflow slots (one workgroup per tile).
__global__ __launch_bounds__(NT) void k_materialize(KParams p, int slot) {
    const InstDesc D = p.desc[slot];
    for (int lt = blockIdx.x; lt < D.ntiles; lt += gridDim.x) {
        const int t = D.tile0 + lt;
        const int tx = lt % D.tiles_x, ty = lt / D.tiles_x;
        const bool matd = p.mat[t] != 0;
        for (int j = threadIdx.x; j < TN; j += NT) {
            const int x = j % TW, y = j / TW;
            const size_t g = (size_t)t * TN + j;
            int4 c = matd ? p.caps[g] : input_caps(D, tx * TW + x, ty * TH + y);
            const int4 f = inflow_of(p, t, x, y);
            c.x += f.x;
            c.y += f.y;
            c.z += f.z;
            c.w += f.w;
            p.caps[g] = c;
        }
        __syncthreads();  // every inflow slot has been read
        for (int k = threadIdx.x; k < INFLOW; k += NT) p.inflow[(size_t)t * INFLOW + k] = 0;
        if (threadIdx.x == 0) p.mat[t] = 1;
        __syncthreads();
    }
}

// Residual state in image layout (certificates and debugging).
__global__ void k_read_state(KParams p, int slot, int32_t *caps_out, int32_t *exc_out, int32_t *dist_out,
                             int32_t *caps8_out) {
    const InstDesc D = p.desc[slot];
    const int64_t n = (int64_t)D.W * D.H;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int X = (int)(i % D.W), Y = (int)(i / D.W);
        const int t = D.tile0 + (Y / TH) * D.tiles_x + X / TW;
        const size_t g = (size_t)t * TN + (Y % TH) * TW + (X % TW);
        int4 c = p.mat[t] ? p.caps[g] : input_caps(D, X, Y);
        const int4 f = inflow_of(p, t, X % TW, Y % TH);
        c.x += f.x;
        c.y += f.y;
        c.z += f.z;
        c.w += f.w;
        if (caps_out) {
            caps_out[4 * i + 0] = c.x;
            caps_out[4 * i + 1] = c.y;
            caps_out[4 * i + 2] = c.z;
            caps_out[4 * i + 3] = c.w;
        }
        if (exc_out) exc_out[i] = p.exc[g];
        if (dist_out) dist_out[i] = p.dist[g];
        if (caps8_out) {
            const int4 c8 = p.caps8[g];
            caps8_out[4 * i + 0] = c8.x;
            caps8_out[4 * i + 1] = c8.y;
            caps8_out[4 * i + 2] = c8.z;
            caps8_out[4 * i + 3] = c8.w;
        }
    }
}

// ---------------------------------------------------------------------------
void launch_arm(const KParams &p, const int32_t *d_slots, int n, int warm, int sweeps, hipStream_t s,
                int clear_tiles) {
    if (n <= 0) return;
    // The wavefront build performs the first global-relabel sweep itself.
    const int start = (!warm && p.variant) ? PH_GR : PH_START;
    const int work = clear_tiles > n ? clear_tiles : n;
    const int grid = (work + 255) / 256 < 65536 ? (work + 255) / 256 : 65536;
    hipLaunchKernelGGL(k_arm, dim3(grid), dim3(256), 0, s, p, d_slots, n, warm, sweeps, start, clear_tiles);
}

template <typename F, typename... Args>
static void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t s, const LaunchEvents &ev, Args... args);

void launch_build(const KParams &p, const int32_t *d_slots, int n, int total_tiles, hipStream_t s,
                  const LaunchEvents &ev) {
    if (n <= 0 || total_tiles <= 0) return;
    if (p.variant) {
        int grid = total_tiles < 65536 ? total_tiles : 65536;
        if (grid >= NUM_XCD) grid -= grid % NUM_XCD;  // whole rounds over the XCDs
        if (p.nbhd == 8)
            launch_timed(k_buildw<8>, dim3(grid), dim3(64), s, ev, p, (const int32_t *)p.bmap,
                         (const TileInfo *)p.tiles, (const InstDesc *)p.desc, total_tiles);
        else
            launch_timed(k_buildw<4>, dim3(grid), dim3(64), s, ev, p, (const int32_t *)p.bmap,
                         (const TileInfo *)p.tiles, (const InstDesc *)p.desc, total_tiles);
        return;
    }
    const int grid = total_tiles < 8192 ? total_tiles : 8192;
    launch_timed(k_build, dim3(grid), dim3(NT), s, ev, p, d_slots, (const int32_t *)(d_slots + n), n, total_tiles);
}

// With events (profile mode) a launch goes through hipExtLaunchKernelGGL,
// whose start/stop events are stamped by the kernel's own dispatch, so the
// measured time is the kernel's execution and not the gap between two queued
// event records.
template <typename F, typename... Args>
static void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t s, const LaunchEvents &ev, Args... args) {
    if (ev.start)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev.start, ev.stop, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

void launch_sched(const KParams &p, int parity, hipStream_t s, const LaunchEvents &ev) {
    if (p.sched_wide)
        launch_timed(k_sched<1024>, dim3(p.n_slots), dim3(1024), s, ev, p, parity);
    else
        launch_timed(k_sched<256>, dim3(p.n_slots), dim3(256), s, ev, p, parity);
}

void launch_pr(const KParams &p, int grid, int parity, hipStream_t s, const LaunchEvents &ev) {
    const int32_t *q = p.queue_pr;
    const TileInfo *ti = p.tiles;
    const InstDesc *de = p.desc;
    const InstState *st = p.state;
    if (p.variant && p.nbhd == 8 && p.inner_iters == 1)
        launch_timed(k_prw<true, 8>, dim3(grid), dim3(64), s, ev, p, parity, q, ti, de, st);
    else if (p.variant && p.nbhd == 8)
        launch_timed(k_prw<false, 8>, dim3(grid), dim3(64), s, ev, p, parity, q, ti, de, st);
    else if (p.variant && p.inner_iters == 1)
        launch_timed(k_prw<true, 4>, dim3(grid), dim3(64), s, ev, p, parity, q, ti, de, st);
    else if (p.variant)
        launch_timed(k_prw<false, 4>, dim3(grid), dim3(64), s, ev, p, parity, q, ti, de, st);
    else if (p.nbhd == 8)
        launch_timed(k_pr8, dim3(grid), dim3(NT), s, ev, p, parity);
    else
        launch_timed(k_pr, dim3(grid), dim3(NT), s, ev, p, parity);
}
// grid: blocks (one-tile blocks: wavefronts, GR_PACK to a workgroup).
template <int BX, int BY, int NB = 4>
static void launch_grb(const KParams &p, int grid, int parity, hipStream_t s, const LaunchEvents &ev) {
    const int wg = BX * BY == 1 ? (grid + GR_PACK - 1) / GR_PACK : grid;
    launch_timed(k_grb<BX, BY, NB>, dim3(wg), dim3(grb_threads<BX, BY>()), s, ev, p, parity,
                 (const int32_t *)p.queue_gr, (const TileInfo *)p.tiles, (const InstDesc *)p.desc,
                 (const InstState *)p.state);
}

int pr_resident_waves_per_cu(const KParams &p) {
    // One value per kernel (the code objects are the same on every device of
    // this architecture); a racing first call computes the same number.
    static std::atomic<int> cache[4];
    const int key = (p.nbhd == 8 ? 2 : 0) + (p.inner_iters == 1 ? 1 : 0);
    int v = cache[key].load(std::memory_order_relaxed);
    if (v > 0) return v;
    const void *k = p.nbhd == 8 ? (p.inner_iters == 1 ? (const void *)k_prw<true, 8> : (const void *)k_prw<false, 8>)
                                : (p.inner_iters == 1 ? (const void *)k_prw<true, 4> : (const void *)k_prw<false, 4>);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64, 0) != hipSuccess || blocks <= 0) {
        (void)hipGetLastError();
        return 16;
    }
    cache[key].store(blocks, std::memory_order_relaxed);
    return blocks;
}

bool gr_block_supported(int bx, int by) {
    switch (bx * 100 + by) {
        case 101: case 102: case 104: case 108: case 116: case 202: case 204: case 208: case 404: return true;
        default: return false;
    }
}

void launch_gr(const KParams &p, int grid, int parity, hipStream_t s, const LaunchEvents &ev) {
    if (!p.variant) {
        if (p.nbhd == 8)
            launch_timed(k_gr8, dim3(grid), dim3(NT), s, ev, p, parity);
        else
            launch_timed(k_gr, dim3(grid), dim3(NT), s, ev, p, parity);
        return;
    }
    if (p.nbhd == 8) {  // one-tile blocks only
        launch_grb<1, 1, 8>(p, grid, parity, s, ev);
        return;
    }
    switch (p.gbx * 100 + p.gby) {
        case 101: launch_grb<1, 1>(p, grid, parity, s, ev); break;
        case 102: launch_grb<1, 2>(p, grid, parity, s, ev); break;
        case 104: launch_grb<1, 4>(p, grid, parity, s, ev); break;
        case 108: launch_grb<1, 8>(p, grid, parity, s, ev); break;
        case 116: launch_grb<1, 16>(p, grid, parity, s, ev); break;
        case 202: launch_grb<2, 2>(p, grid, parity, s, ev); break;
        case 204: launch_grb<2, 4>(p, grid, parity, s, ev); break;
        case 208: launch_grb<2, 8>(p, grid, parity, s, ev); break;
        case 404: launch_grb<4, 4>(p, grid, parity, s, ev); break;
        default: break;  // rejected when the options are set
    }
}

void launch_step(const KParams &p, int grid_pr, int grid_gr, int parity, hipStream_t s) {
    launch_sched(p, parity, s, LaunchEvents{});
    launch_pr(p, grid_pr, parity, s, LaunchEvents{});
    launch_gr(p, grid_gr, parity, s, LaunchEvents{});
}


void launch_edit(const KParams &p, int slot, const int32_t *d_strokes, int n_strokes, int32_t delta,
                 hipStream_t s) {
    if (n_strokes <= 0) return;
    hipLaunchKernelGGL(k_edit, dim3(n_strokes < 1024 ? n_strokes : 1024), dim3(NT), 0, s, p, slot, d_strokes,
                       n_strokes, delta);
    hipLaunchKernelGGL(k_edit_masks, dim3(n_strokes < 1024 ? n_strokes : 1024), dim3(NT), 0, s, p, slot, d_strokes,
                       n_strokes);
}

void launch_materialize(const KParams &p, int slot, int ntiles, hipStream_t s) {
    if (ntiles <= 0) return;
    hipLaunchKernelGGL(k_materialize, dim3(ntiles < 4096 ? ntiles : 4096), dim3(NT), 0, s, p, slot);
}

void launch_reframe(const KParams &p, int slot, int ntiles, const int32_t *const *in1, int32_t pen1,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_reframe, dim3(1024), dim3(NT), 0, s, p, slot, in1[0], in1[1], in1[2], in1[3], in1[4],
                       in1[5], pen1);
    hipLaunchKernelGGL(k_remask, dim3(ntiles < 4096 ? ntiles : 4096), dim3(NT), 0, s, p, slot);
}

void launch_read_state(const KParams &p, int slot, int32_t *d_caps, int32_t *d_exc, int32_t *d_dist,
                       int32_t *d_caps8, hipStream_t s) {
    hipLaunchKernelGGL(k_read_state, dim3(1024), dim3(256), 0, s, p, slot, d_caps, d_exc, d_dist, d_caps8);
}

}  // namespace gcfg
