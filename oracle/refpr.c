/*
 * CPU restatement of the reference push-relabel schedule -- TEST
 * INFRASTRUCTURE ONLY (also timed as the CPU baseline; never linked into the
 * product).
 *
 * It follows GraphCut.cu:385-534 (GC_Optimize) driving the kernels of
 * GraphCutKernels.cu with the knobs of GraphCut.h:22-171 (NEIGHBORHOOD 4,
 * SPREAD_ZEROS, 32x8 blocks, 40 pushes / 10 relabels / 10 GR loops per launch,
 * activity check every 10 iterations, global relabel at iterations 100, 200...).
 *
 * Each CUDA launch is executed as blocks in a 4-colour order (bx%2, by%2);
 * blocks of one colour have disjoint footprints (own 32x8 cells plus a
 * one-cell halo), so this is one legal CUDA schedule and the colour classes
 * run in parallel under OpenMP. Inside a block, every phase between two
 * __syncthreads is executed as "all threads read, then all threads write",
 * which is also a legal interleaving of the warps.
 */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define TX 32
#define TY 8
#define LW (TX + 2)
#define LSZ ((TX + 2) * (TY + 2))
#define PUSHES_PER_KERNEL 40
#define RELABELS_PER_KERNEL 10
#define GR_LOOPS 10
#define ACTIVITY_CHECK_FREQUENCY 10
#define GLOBAL_RELABEL_FREQUENCY 100
#define FIRST_GLOBAL_RELABEL 100

typedef struct {
    int w, h, size, wex, hex, sizeex, bx, by;
    int32_t *el, *er, *eu, *ed, *height, *excess, *status, *compn, *active;
} rg_t;

static inline int interior(const rg_t *g, int x, int y) { return x > 0 && y > 0 && x < g->w - 1 && y < g->h - 1; }

/* GraphCutKernels.cu:165-214 (InitGraphVarEdges) and :112-157 (InitGraph). */
static void init_graph(rg_t *g, const int32_t *pos, const int32_t *neg, const int32_t *up, const int32_t *down,
                       const int32_t *left, const int32_t *right, int32_t penalty) {
    for (int b = 0; b < g->bx * g->by; ++b) g->active[b] = 0;
    for (int y = 0; y < g->hex; ++y)
        for (int x = 0; x < g->wex; ++x) {
            const int t = x + y * g->wex;
            const int in = interior(g, x, y);
            const int blk = x / TX + (y / TY) * g->bx;
            if (in) g->active[blk] = 1;
            const int i = x + y * g->w;
            const int coming = in ? pos[i] : 0, going = in ? neg[i] : 0;
            g->eu[t] = in ? (up ? up[i] : penalty) : 0;
            g->ed[t] = in ? (up ? down[i] : penalty) : 0;
            g->el[t] = in ? (up ? left[i] : penalty) : 0;
            g->er[t] = in ? (up ? right[i] : penalty) : 0;
            g->height[t] = going >= coming ? 1 : 2;
            g->excess[t] = coming - going;
            g->status[t] = in ? (going >= coming ? 2 : 1) : 0;
            if (x == 1) g->el[t] = 0;
            if (x == g->w - 2) g->er[t] = 0;
            if (y == 1) g->eu[t] = 0;
            if (y == g->h - 2) g->ed[t] = 0;
            g->compn[t] = 0;
        }
}

/* Load own heights plus the one-cell halo with the bounds of
 * GraphCutKernels.cu:616-629 (cells that the reference leaves unloaded are
 * never read through a non-zero edge, so they are filled with 0 here). */
static void load_heights(const rg_t *g, int bx, int by, int32_t *lh) {
    memset(lh, 0, sizeof(int32_t) * LSZ);
    for (int ty = 0; ty < TY; ++ty)
        for (int tx = 0; tx < TX; ++tx) {
            const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex;
            const int li = (ty + 1) * LW + tx + 1;
            lh[li] = g->height[t];
            if (tx == TX - 1 && x < g->wex - 1) lh[li + 1] = g->height[t + 1];
            if (tx == 0 && x > 0) lh[li - 1] = g->height[t - 1];
            if (ty == TY - 1 && y < g->hex - 1) lh[li + LW] = g->height[t + g->wex];
            if (ty == 0 && y > 0) lh[li - LW] = g->height[t - g->wex];
        }
}

/* GraphCutKernels.cu:600-724 (Relabel) for one block. */
static void relabel_block(rg_t *g, int bx, int by, int skip) {
    if (skip && !g->active[bx + by * g->bx]) return;
    int32_t lh[LSZ], hnew[TX * TY], hres[TX * TY];
    load_heights(g, bx, by, lh);
    for (int k = 0; k < TX * TY; ++k) hres[k] = -1;
    for (int it = 0; it < RELABELS_PER_KERNEL; ++it) {
        for (int ty = 0; ty < TY; ++ty)
            for (int tx = 0; tx < TX; ++tx) {
                const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex;
                const int li = (ty + 1) * LW + tx + 1, k = ty * TX + tx;
                hnew[k] = -2;
                if (g->excess[t] >= 0 && g->status[t] > 0) {
                    int hh = g->size;
                    if (hh > lh[li - 1] && g->el[t] > 0) hh = lh[li - 1];
                    if (hh > lh[li + 1] && g->er[t] > 0) hh = lh[li + 1];
                    if (hh > lh[li - LW] && g->eu[t] > 0) hh = lh[li - LW];
                    if (hh > lh[li + LW] && g->ed[t] > 0) hh = lh[li + LW];
                    hnew[k] = hh;
                    hres[k] = hh;
                }
            }
        for (int ty = 0; ty < TY; ++ty)
            for (int tx = 0; tx < TX; ++tx) {
                const int k = ty * TX + tx;
                if (hnew[k] != -2) lh[(ty + 1) * LW + tx + 1] = hnew[k] + 1;
            }
    }
    for (int ty = 0; ty < TY; ++ty)
        for (int tx = 0; tx < TX; ++tx) {
            const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex, k = ty * TX + tx;
            const int hh = hres[k];
            if (hh != -1) g->height[t] = hh + 1;
            if (g->status[t]) g->status[t] = (hh == g->size) ? 0 : 2 - (g->excess[t] > 0);
        }
}

/* GraphCutKernels.cu:433-581 (WavePush with DO_PUSH_C_WAVE2, :83-99) for one block. */
static void push_block(rg_t *g, int bx, int by) {
    if (!g->active[bx + by * g->bx]) return;
    int32_t lh[LSZ], lex[LSZ], snap[TX * TY], orig[TX * TY];
    load_heights(g, bx, by, lh);
    memset(lex, 0, sizeof(lex));
    for (int ty = 0; ty < TY; ++ty)
        for (int tx = 0; tx < TX; ++tx) {
            const int x = bx * TX + tx, y = by * TY + ty;
            const int li = (ty + 1) * LW + tx + 1;
            lex[li] = g->excess[x + y * g->wex];
            orig[ty * TX + tx] = lex[li];
        }
    int32_t *edge[4] = {g->el, g->er, g->eu, g->ed};
    int32_t *inv[4] = {g->er, g->el, g->ed, g->eu};
    const int lgap[4] = {-1, 1, -LW, LW};
    const int ggap[4] = {-1, 1, -g->wex, g->wex};
    for (int i = 0; i < PUSHES_PER_KERNEL; ++i) {
        int did = 0;
        for (int dir = 0; dir < 4; ++dir) {
            for (int ty = 0; ty < TY; ++ty)
                for (int tx = 0; tx < TX; ++tx) snap[ty * TX + tx] = lex[(ty + 1) * LW + tx + 1];
            for (int ty = 0; ty < TY; ++ty)
                for (int tx = 0; tx < TX; ++tx) {
                    const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex;
                    const int li = (ty + 1) * LW + tx + 1;
                    const int ex = snap[ty * TX + tx];
                    const int cap = edge[dir][t];
                    if (cap > 0 && ex > 0 && lh[li] == lh[li + lgap[dir]] + 1) {
                        const int f = ex > cap ? cap : ex;
                        edge[dir][t] -= f;
                        inv[dir][t + ggap[dir]] += f;
                        lex[li] -= f;
                        lex[li + lgap[dir]] += f;
                        did = 1;
                    }
                }
        }
        if (!did) break;
    }
    for (int ty = 0; ty < TY; ++ty)
        for (int tx = 0; tx < TX; ++tx) {
            const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex;
            const int li = (ty + 1) * LW + tx + 1;
            const int d = lex[li] - orig[ty * TX + tx];
            if (d) g->excess[t] += d;
            if (tx == 0 && x > 0 && lex[li - 1]) g->excess[t - 1] += lex[li - 1];
            if (ty == 0 && y > 0 && lex[li - LW]) g->excess[t - g->wex] += lex[li - LW];
            if (tx == TX - 1 && x < g->wex - 1 && lex[li + 1]) g->excess[t + 1] += lex[li + 1];
            if (ty == TY - 1 && y < g->hex - 1 && lex[li + LW]) g->excess[t + g->wex] += lex[li + LW];
        }
}

/* GraphCutKernels.cu:757-783 (InitGlobalRelabel). */
static void init_global_relabel(rg_t *g) {
    for (int y = 0; y < g->hex; ++y)
        for (int x = 0; x < g->wex; ++x) {
            const int t = x + y * g->wex;
            int no_sink = 0;
            if (interior(g, x, y)) {
                no_sink = g->excess[t] >= 0;
                g->height[t] = no_sink * g->size + 1;
            }
            int c = (g->el[t] > 0) | ((g->er[t] > 0) << 1) | ((g->eu[t] > 0) << 2) | ((g->ed[t] > 0) << 3);
            g->compn[t] = c | (no_sink << 8);
        }
}

/* GraphCutKernels.cu:796-875 (GlobalRelabel) for one block; returns alive. */
static int global_relabel_block(rg_t *g, int bx, int by) {
    int any = 0;
    for (int ty = 0; ty < TY && !any; ++ty)
        for (int tx = 0; tx < TX; ++tx)
            if (g->compn[(bx * TX + tx) + (by * TY + ty) * g->wex] & 256) {
                any = 1;
                break;
            }
    if (!any) return 0;
    int32_t lh[LSZ], hn[TX * TY];
    uint8_t changed[TX * TY];
    load_heights(g, bx, by, lh);
    memset(changed, 0, sizeof(changed));
    for (int rep = 0; rep < GR_LOOPS; ++rep) {
        for (int ty = 0; ty < TY; ++ty)
            for (int tx = 0; tx < TX; ++tx) {
                const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex;
                const int li = (ty + 1) * LW + tx + 1, cn = g->compn[t];
                int hh = lh[li] - 1;
                if ((cn & 256) && interior(g, x, y)) {
                    hh = g->size;
                    if (hh > lh[li - 1] && (cn & 1)) hh = lh[li - 1];
                    if (hh > lh[li + 1] && (cn & 2)) hh = lh[li + 1];
                    if (hh > lh[li - LW] && (cn & 4)) hh = lh[li - LW];
                    if (hh > lh[li + LW] && (cn & 8)) hh = lh[li + LW];
                }
                hn[ty * TX + tx] = hh;
            }
        for (int ty = 0; ty < TY; ++ty)
            for (int tx = 0; tx < TX; ++tx) {
                const int li = (ty + 1) * LW + tx + 1, k = ty * TX + tx;
                changed[k] |= lh[li] != hn[k] + 1;
                lh[li] = hn[k] + 1;
            }
    }
    int alive = 0;
    for (int ty = 0; ty < TY; ++ty)
        for (int tx = 0; tx < TX; ++tx) {
            const int x = bx * TX + tx, y = by * TY + ty, t = x + y * g->wex;
            if (changed[ty * TX + tx]) {
                g->compn[t] ^= 256;
                g->height[t] = lh[(ty + 1) * LW + tx + 1];
                alive = 1;
            }
        }
    return alive;
}

/* GraphCutKernels.cu:729-743 (UpdateActivity); returns alive. */
static int update_activity(rg_t *g) {
    int alive = 0;
    for (int b = 0; b < g->bx * g->by; ++b) g->active[b] = 0;
    for (int y = 0; y < g->hex; ++y)
        for (int x = 0; x < g->wex; ++x)
            if (g->status[x + y * g->wex] == 1) {
                g->active[x / TX + (y / TY) * g->bx] = 1;
                alive = 1;
            }
    return alive;
}

/* GraphCutKernels.cu:228-258 (InitLabels) and :265-351 (SpreadLabels). */
static void labels(rg_t *g, int32_t *label) {
    const int w = g->w, h = g->h;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int t = x + y * g->wex;
            label[x + y * w] = interior(g, x, y) ? g->excess[t] >= 0 : 0;
        }
    for (int y = 0; y < g->hex; ++y)
        for (int x = 0; x < g->wex; ++x) {
            const int t = x + y * g->wex;
            g->compn[t] = (!g->ed[t]) | ((!g->eu[t]) << 1) | ((!g->er[t]) << 2) | ((!g->el[t]) << 3);
        }
    int alive = 1;
    while (alive) {
        alive = 0;
        for (int color = 0; color < 4; ++color) {
#pragma omp parallel for schedule(dynamic, 4) reduction(| : alive)
            for (int b = 0; b < g->bx * g->by; ++b) {
                const int bx = b % g->bx, by = b / g->bx;
                if (((bx & 1) | ((by & 1) << 1)) != color) continue;
                int32_t ll[LSZ], cur[TX * TY], nxt[TX * TY];
                memset(ll, 0, sizeof(ll));
                for (int ty = 0; ty < TY; ++ty)
                    for (int tx = 0; tx < TX; ++tx) {
                        const int x = bx * TX + tx, y = by * TY + ty, li = (ty + 1) * LW + tx + 1;
                        if (x < w && y < h) {
                            const int i = x + y * w;
                            ll[li] = label[i];
                            if (tx == TX - 1 && x < w - 1) ll[li + 1] = label[i + 1];
                            if (tx == 0 && x > 0) ll[li - 1] = label[i - 1];
                            if (ty == TY - 1 && y < h - 1) ll[li + LW] = label[i + w];
                            if (ty == 0 && y > 0) ll[li - LW] = label[i - w];
                        }
                        cur[ty * TX + tx] = ll[li];
                    }
                for (int rep = 0; rep < GR_LOOPS; ++rep) {
                    for (int ty = 0; ty < TY; ++ty)
                        for (int tx = 0; tx < TX; ++tx) {
                            const int x = bx * TX + tx, y = by * TY + ty, li = (ty + 1) * LW + tx + 1;
                            const int k = ty * TX + tx;
                            int c = cur[k];
                            nxt[k] = c;
                            if (c && interior(g, x, y)) {
                                const int cn = g->compn[x + y * g->wex];
                                c = ((cn & 1) || ll[li + LW]) && c;
                                c = ((cn & 2) || ll[li - LW]) && c;
                                c = ((cn & 4) || ll[li + 1]) && c;
                                c = ((cn & 8) || ll[li - 1]) && c;
                                nxt[k] = c;
                            }
                        }
                    for (int ty = 0; ty < TY; ++ty)
                        for (int tx = 0; tx < TX; ++tx) {
                            const int x = bx * TX + tx, y = by * TY + ty, k = ty * TX + tx;
                            if (nxt[k] != cur[k]) {
                                label[x + y * w] = nxt[k];
                                ll[(ty + 1) * LW + tx + 1] = nxt[k];
                                alive = 1;
                                cur[k] = nxt[k];
                            }
                        }
                }
            }
        }
    }
}

static void launch_colour(rg_t *g, int kernel, int skip, int *alive_out) {
    int alive = 0;
    for (int color = 0; color < 4; ++color) {
#pragma omp parallel for schedule(dynamic, 4) reduction(| : alive)
        for (int b = 0; b < g->bx * g->by; ++b) {
            const int bx = b % g->bx, by = b / g->bx;
            if (((bx & 1) | ((by & 1) << 1)) != color) continue;
            if (kernel == 0)
                relabel_block(g, bx, by, skip);
            else if (kernel == 1)
                push_block(g, bx, by);
            else
                alive |= global_relabel_block(g, bx, by);
        }
    }
    if (alive_out) *alive_out = alive;
}

int oracle_refpr_solve(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                       const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                       int threads, int32_t *label, int64_t *flow, int64_t *printed_energy) {
    if (width <= 0 || height <= 0 || !pos || !neg || !label || !flow) return -1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    rg_t g;
    memset(&g, 0, sizeof(g));
    g.w = width;
    g.h = height;
    g.size = width * height;
    g.wex = (width + TX - 1) / TX * TX;
    g.hex = (height + TY - 1) / TY * TY;
    g.sizeex = g.wex * g.hex;
    g.bx = g.wex / TX;
    g.by = g.hex / TY;
    int32_t **planes[9] = {&g.el, &g.er, &g.eu, &g.ed, &g.height, &g.excess, &g.status, &g.compn, &g.active};
    for (int p = 0; p < 9; ++p) {
        *planes[p] = (int32_t *)calloc((size_t)g.sizeex, sizeof(int32_t));
        if (!*planes[p]) return -2;
    }
    init_graph(&g, pos, neg, up, down, left, right, penalty);

    /* Printed "Initial energy level" (GraphCut.cu:399-401): limited-mode
     * EnergyLevel sums indices < W*H of a W_ex-strided array (:887-897). */
    int64_t initial = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const int64_t k = x + (int64_t)y * g.wex;
            if (k < (int64_t)width * height) initial += pos[x + y * width] > 0 ? pos[x + y * width] : 0;
        }

    int counter = 0, h_alive = 1, iters = 0;
    const int max_iters = 2000000;
    for (;;) {
        const int skip = counter % ACTIVITY_CHECK_FREQUENCY;
        ++counter;
        launch_colour(&g, 0, skip, NULL);
        launch_colour(&g, 1, skip, NULL);
        if (!skip) h_alive = update_activity(&g);
        if (!h_alive) break;
        if (!((counter - FIRST_GLOBAL_RELABEL) % GLOBAL_RELABEL_FREQUENCY)) {
            init_global_relabel(&g);
            int alive = 1;
            while (alive) launch_colour(&g, 2, 0, &alive);
            h_alive = 1;
        }
        if (++iters > max_iters) return -5;
    }
    int64_t stuck = 0, stuck_all = 0, sp = 0;
    for (int y = 0; y < g.hex; ++y)
        for (int x = 0; x < g.wex; ++x) {
            const int e = g.excess[x + y * g.wex];
            if (e > 0) stuck_all += e;
            if (interior(&g, x, y)) {
                if (e > 0) stuck += e;
                sp += pos[x + y * width];
            }
        }
    *flow = sp - stuck;
    if (printed_energy) *printed_energy = initial - stuck_all;
    labels(&g, label);
    for (int p = 0; p < 9; ++p) free(*planes[p]);
    return counter;
}
