/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h). Never linked into
 * the product library.
 *
 * oracle_maxflow_grid: FIFO push-relabel with exact global relabelling,
 * int64 residuals and excess, on the graph of GraphCutKernels.cu:165-214
 * (interior nodes only; arcs into the border zeroed at :200-203). The
 * returned flow is the exact max-flow value over the interior; labels are the
 * set of interior nodes that cannot reach the sink in the final residual graph
 * (SPREAD_ZEROS semantics of GraphCutKernels.cu:237-238,315-320), which is the
 * same set for every maximum flow.
 *
 * oracle_dataterm: direct per-pixel evaluation of
 *   p_k(I) = prod_c (n_{k,c}[I_c >> 4] + 1) / (N_k + 16)
 *   pos = llround(64 * -ln p_bg(I)),  neg = llround(64 * -ln p_fg(I))
 *   beta = (sum over 4-neighbour pairs of |I_p - I_q|^2) / #pairs
 *   w = llround(lambda * exp(-|I_p - I_q|^2 / (2 beta)))   (w = lambda if beta == 0)
 * This colour stage replaces the reference's 256^3 Gaussian KDE
 * (DataTermKernels.cu:61-83,116-150) behind the same DataTerm.h API shape.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DINF 0x3fffffff

static inline int is_interior(int x, int y, int w, int h) { return x > 0 && y > 0 && x < w - 1 && y < h - 1; }

/* Directions: 0 left, 1 right, 2 up, 3 down, then the diagonals of the
 * 8-neighbourhood (GraphCut.h:65-70): 4 up-left, 5 up-right, 6 down-left,
 * 7 down-right. OPP is the reverse arc. */
static const int DX[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
static const int DY[8] = {0, 0, -1, 1, -1, -1, 1, 1};
static const int OPP[8] = {1, 0, 3, 2, 7, 6, 5, 4};

/* Capacity of the arc from (x,y) in direction dir, zero unless both ends are
 * interior (GraphCutKernels.cu:143-156,200-213). planes[dir] == NULL for all
 * directions means every n-link is penalty (GraphCutKernels.cu:129-140). */
static int64_t arc_cap(int x, int y, int w, int h, int dir, const int32_t *const *planes, int32_t penalty) {
    if (!is_interior(x, y, w, h) || !is_interior(x + DX[dir], y + DY[dir], w, h)) return 0;
    if (!planes[0]) return penalty;
    return planes[dir][(size_t)y * w + x];
}

typedef struct {
    int w, h, n_interior, ndir;
    size_t n;
    long off[8];
    int64_t *e, *rt, *r;
    int32_t *d, *queue;
    uint8_t *inq;
    size_t qhead, qtail, qcap;
} pr_t;

static void q_push(pr_t *g, int32_t v) {
    g->queue[g->qtail] = v;
    g->qtail = (g->qtail + 1) % g->qcap;
    g->inq[v] = 1;
}

static int32_t q_pop(pr_t *g) {
    int32_t v = g->queue[g->qhead];
    g->qhead = (g->qhead + 1) % g->qcap;
    g->inq[v] = 0;
    return v;
}

/* Exact BFS distances to the sink over residual arcs (nodes with sink residual
 * are at distance 1); unreachable nodes get DINF. Rebuilds the active queue. */
static void global_relabel(pr_t *g) {
    const int w = g->w, h = g->h, K = g->ndir;
    int32_t *bfs = (int32_t *)malloc(sizeof(int32_t) * (g->n ? g->n : 1));
    size_t head = 0, tail = 0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            g->d[i] = DINF;
            if (is_interior(x, y, w, h) && g->rt[i] > 0) {
                g->d[i] = 1;
                bfs[tail++] = (int32_t)i;
            }
        }
    while (head < tail) {
        const int32_t v = bfs[head++];
        for (int dir = 0; dir < K; ++dir) {
            const long u = (long)v + g->off[dir];
            if (u < 0 || (size_t)u >= g->n) continue;
            if (g->d[u] != DINF) continue;
            if (g->r[(size_t)u * K + OPP[dir]] <= 0) continue;
            g->d[u] = g->d[v] + 1;
            bfs[tail++] = (int32_t)u;
        }
    }
    free(bfs);
    g->qhead = g->qtail = 0;
    memset(g->inq, 0, g->n);
    for (size_t i = 0; i < g->n; ++i)
        if (g->e[i] > 0 && g->d[i] < DINF) q_push(g, (int32_t)i);
}

static int maxflow_core(int width, int height, const int32_t *pos, const int32_t *neg,
                        const int32_t *const *planes, int ndir, int32_t penalty, int32_t *label, int64_t *flow_out,
                        int32_t *caps_out, int32_t *exc_out) {
    if (width <= 0 || height <= 0 || !pos || !neg || !label || !flow_out) return -1;
    if (planes[0])
        for (int k = 1; k < ndir; ++k)
            if (!planes[k]) return -1;
    const size_t n = (size_t)width * height;
    memset(label, 0, sizeof(int32_t) * n);
    *flow_out = 0;
    if (width < 3 || height < 3) return 0;
    const int K = ndir;

    pr_t g;
    memset(&g, 0, sizeof(g));
    g.w = width;
    g.h = height;
    g.n = n;
    g.ndir = K;
    g.n_interior = (width - 2) * (height - 2);
    for (int dir = 0; dir < K; ++dir) g.off[dir] = DX[dir] + (long)DY[dir] * width;
    g.e = (int64_t *)calloc(n, sizeof(int64_t));
    g.rt = (int64_t *)calloc(n, sizeof(int64_t));
    g.r = (int64_t *)calloc(n * K, sizeof(int64_t));
    g.d = (int32_t *)calloc(n, sizeof(int32_t));
    g.qcap = n + 1;
    g.queue = (int32_t *)calloc(g.qcap, sizeof(int32_t));
    g.inq = (uint8_t *)calloc(n, 1);
    if (!g.e || !g.rt || !g.r || !g.d || !g.queue || !g.inq) return -2;

    int64_t flow = 0;
    int bad = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            if (!is_interior(x, y, width, height)) continue;
            for (int dir = 0; dir < K; ++dir) {
                const int64_t c = arc_cap(x, y, width, height, dir, planes, penalty);
                if (c < 0) bad = 1;
                g.r[i * K + dir] = c;
            }
            /* t-links may be negative, as in the reference (GraphCutKernels.cu:196-197):
             * subtracting min(p, q) from both is a reparametrisation, and the
             * returned value is the minimum cut energy with the t-links as given. */
            const int64_t p = pos[i], q = neg[i];
            const int64_t m = p < q ? p : q;
            flow += m;
            g.e[i] = p - m;
            g.rt[i] = q - m;
        }
    if (bad) {
        free(g.e), free(g.rt), free(g.r), free(g.d), free(g.queue), free(g.inq);
        return -3;
    }

    global_relabel(&g);
    long relabels = 0;
    const long gr_every = g.n_interior / 2 + 16;
    while (g.qhead != g.qtail) {
        const int32_t u = q_pop(&g);
        while (g.e[u] > 0 && g.d[u] < DINF) {
            if (g.rt[u] > 0) {
                const int64_t dlt = g.e[u] < g.rt[u] ? g.e[u] : g.rt[u];
                g.e[u] -= dlt;
                g.rt[u] -= dlt;
                flow += dlt;
                continue;
            }
            for (int dir = 0; dir < K && g.e[u] > 0; ++dir) {
                const int64_t rr = g.r[(size_t)u * K + dir];
                if (rr <= 0) continue;
                const long v = (long)u + g.off[dir];
                if (g.d[u] != g.d[v] + 1) continue;
                const int64_t dlt = g.e[u] < rr ? g.e[u] : rr;
                g.r[(size_t)u * K + dir] -= dlt;
                g.r[(size_t)v * K + OPP[dir]] += dlt;
                g.e[u] -= dlt;
                g.e[v] += dlt;
                if (!g.inq[v] && g.d[v] < DINF) q_push(&g, (int32_t)v);
            }
            if (g.e[u] > 0) {
                int32_t m = DINF;
                for (int dir = 0; dir < K; ++dir) {
                    if (g.r[(size_t)u * K + dir] <= 0) continue;
                    const int32_t dv = g.d[(long)u + g.off[dir]];
                    if (dv < m) m = dv;
                }
                g.d[u] = (m >= DINF - 1 || m + 1 > g.n_interior) ? DINF : m + 1;
                if (++relabels >= gr_every) {
                    relabels = 0;
                    if (g.e[u] > 0 && g.d[u] < DINF && !g.inq[u]) q_push(&g, u);
                    global_relabel(&g);
                    break;
                }
            }
        }
    }
    global_relabel(&g);
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            label[i] = is_interior(x, y, width, height) && g.d[i] == DINF;
        }
    /* Consistency: flow == sum_interior pos - sum stuck excess. */
    int64_t sp = 0, se = 0;
    for (int y = 1; y < height - 1; ++y)
        for (int x = 1; x < width - 1; ++x) {
            const size_t i = (size_t)y * width + x;
            sp += pos[i];
            se += g.e[i] > 0 ? g.e[i] : 0;
        }
    const int ok = (sp - se == flow);
    /* Residual state in the solver's form: excess = e - remaining sink
     * capacity (a node with e > 0 has none left), residual caps per arc. */
    if (caps_out)
        for (size_t i = 0; i < n * (size_t)K; ++i) caps_out[i] = (int32_t)g.r[i];
    if (exc_out)
        for (size_t i = 0; i < n; ++i) exc_out[i] = (int32_t)(g.e[i] - g.rt[i]);
    free(g.e), free(g.rt), free(g.r), free(g.d), free(g.queue), free(g.inq);
    *flow_out = flow;
    return ok ? 0 : -4;
}

int oracle_maxflow_grid(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                        const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                        int32_t *label, int64_t *flow_out) {
    const int32_t *planes[8] = {left, right, up, down, NULL, NULL, NULL, NULL};
    if (up && (!down || !left || !right)) return -1;
    if (!up) planes[0] = planes[1] = planes[2] = planes[3] = NULL;
    return maxflow_core(width, height, pos, neg, planes, 4, penalty, label, flow_out, NULL, NULL);
}

int oracle_maxflow_grid_state(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                              const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                              int32_t *label, int64_t *flow_out, int32_t *caps_out, int32_t *exc_out) {
    const int32_t *planes[8] = {left, right, up, down, NULL, NULL, NULL, NULL};
    if (up && (!down || !left || !right)) return -1;
    if (!up) planes[0] = planes[1] = planes[2] = planes[3] = NULL;
    if (caps_out) memset(caps_out, 0, sizeof(int32_t) * 4 * (size_t)width * height);
    if (exc_out) memset(exc_out, 0, sizeof(int32_t) * (size_t)width * height);
    return maxflow_core(width, height, pos, neg, planes, 4, penalty, label, flow_out, caps_out, exc_out);
}

int oracle_maxflow_grid8(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *const *dirs8,
                         int32_t penalty, int32_t *label, int64_t *flow_out) {
    const int32_t *planes[8] = {NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
    if (dirs8)
        for (int k = 0; k < 8; ++k) planes[k] = dirs8[k];
    return maxflow_core(width, height, pos, neg, planes, 8, penalty, label, flow_out, NULL, NULL);
}

static int64_t cut_core(int width, int height, const int32_t *pos, const int32_t *neg,
                        const int32_t *const *planes, int ndir, int32_t penalty, const int32_t *label) {
    int64_t cut = 0;
    for (int y = 1; y < height - 1; ++y)
        for (int x = 1; x < width - 1; ++x) {
            const size_t i = (size_t)y * width + x;
            if (label[i]) {
                cut += neg[i];
                for (int dir = 0; dir < ndir; ++dir) {
                    const int xx = x + DX[dir], yy = y + DY[dir];
                    if (!is_interior(xx, yy, width, height)) continue;
                    if (label[(size_t)yy * width + xx]) continue;
                    cut += arc_cap(x, y, width, height, dir, planes, penalty);
                }
            } else {
                cut += pos[i];
            }
        }
    return cut;
}

int64_t oracle_cut_value(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                         const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                         const int32_t *label) {
    const int32_t *planes[8] = {up ? left : NULL, right, up, down, NULL, NULL, NULL, NULL};
    return cut_core(width, height, pos, neg, planes, 4, penalty, label);
}

int64_t oracle_cut_value8(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *const *dirs8,
                          int32_t penalty, const int32_t *label) {
    const int32_t *planes[8] = {NULL, NULL, NULL, NULL, NULL, NULL, NULL, NULL};
    if (dirs8)
        for (int k = 0; k < 8; ++k) planes[k] = dirs8[k];
    return cut_core(width, height, pos, neg, planes, 8, penalty, label);
}

static int32_t q_round(double v) { return (int32_t)llround(v); }

int oracle_dataterm(int width, int height, const uint8_t *rgb, const uint8_t *seeds, int32_t lambda,
                    int32_t *hist_out, int32_t *counts_out, int64_t *beta_sum_out, int64_t *pairs_out,
                    double *beta_out, int32_t *pos, int32_t *neg, int32_t *up, int32_t *down, int32_t *left,
                    int32_t *right, double *tpos_f, double *tneg_f, double *right_f, double *down_f) {
    return oracle_dataterm_model(width, height, rgb, rgb, seeds, lambda, hist_out, counts_out, beta_sum_out,
                                 pairs_out, beta_out, pos, neg, up, down, left, right, tpos_f, tneg_f, right_f,
                                 down_f);
}

int oracle_dataterm_model(int width, int height, const uint8_t *rgb, const uint8_t *model_rgb,
                          const uint8_t *seeds, int32_t lambda, int32_t *hist_out, int32_t *counts_out,
                          int64_t *beta_sum_out, int64_t *pairs_out, double *beta_out, int32_t *pos, int32_t *neg,
                          int32_t *up, int32_t *down, int32_t *left, int32_t *right, double *tpos_f,
                          double *tneg_f, double *right_f, double *down_f) {
    if (width <= 0 || height <= 0 || !rgb || !model_rgb || !seeds) return -1;
    int32_t hist[2][3][16];
    int32_t counts[2] = {0, 0};
    memset(hist, 0, sizeof(hist));
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; ++i) {
        const int s = seeds[i];
        if (s != 1 && s != 2) continue;
        const int k = s - 1;
        counts[k]++;
        for (int c = 0; c < 3; ++c) hist[k][c][model_rgb[3 * i + c] >> 4]++;
    }
    int64_t bsum = 0, npairs = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const uint8_t *a = rgb + 3 * ((size_t)y * width + x);
            if (x + 1 < width) {
                const uint8_t *b = a + 3;
                for (int c = 0; c < 3; ++c) bsum += (int64_t)(a[c] - b[c]) * (a[c] - b[c]);
                npairs++;
            }
            if (y + 1 < height) {
                const uint8_t *b = a + 3 * (size_t)width;
                for (int c = 0; c < 3; ++c) bsum += (int64_t)(a[c] - b[c]) * (a[c] - b[c]);
                npairs++;
            }
        }
    const double beta = npairs ? (double)bsum / (double)npairs : 0.0;
    if (hist_out) memcpy(hist_out, hist, sizeof(hist));
    if (counts_out) counts_out[0] = counts[0], counts_out[1] = counts[1];
    if (beta_sum_out) *beta_sum_out = bsum;
    if (pairs_out) *pairs_out = npairs;
    if (beta_out) *beta_out = beta;

    int32_t *wr_q = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *wd_q = (int32_t *)calloc(n, sizeof(int32_t));
    if (!wr_q || !wd_q) {
        free(wr_q), free(wd_q);
        return -2;
    }
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            const uint8_t *a = rgb + 3 * i;
            double cost[2];
            for (int k = 0; k < 2; ++k) {
                double p = 1.0;
                for (int c = 0; c < 3; ++c) p *= (double)(hist[k][c][a[c] >> 4] + 1) / (double)(counts[k] + 16);
                cost[k] = 64.0 * -log(p);
            }
            /* pos is the cost of the background label (source side = foreground). */
            if (tpos_f) tpos_f[i] = cost[1];
            if (tneg_f) tneg_f[i] = cost[0];
            if (pos) pos[i] = q_round(cost[1]);
            if (neg) neg[i] = q_round(cost[0]);
            double wr = 0.0, wd = 0.0;
            if (x + 1 < width) {
                const uint8_t *b = a + 3;
                int s = 0;
                for (int c = 0; c < 3; ++c) s += (a[c] - b[c]) * (a[c] - b[c]);
                wr = beta > 0.0 ? (double)lambda * exp(-(double)s / (2.0 * beta)) : (double)lambda;
                wr_q[i] = q_round(wr);
            }
            if (y + 1 < height) {
                const uint8_t *b = a + 3 * (size_t)width;
                int s = 0;
                for (int c = 0; c < 3; ++c) s += (a[c] - b[c]) * (a[c] - b[c]);
                wd = beta > 0.0 ? (double)lambda * exp(-(double)s / (2.0 * beta)) : (double)lambda;
                wd_q[i] = q_round(wd);
            }
            if (right_f) right_f[i] = wr;
            if (down_f) down_f[i] = wd;
        }
    /* Symmetric weights: left/up are the right/down weight of the neighbour. */
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            if (right) right[i] = wr_q[i];
            if (down) down[i] = wd_q[i];
            if (left) left[i] = x > 0 ? wr_q[i - 1] : 0;
            if (up) up[i] = y > 0 ? wd_q[i - width] : 0;
        }
    free(wr_q);
    free(wd_q);
    return 0;
}

/* Kernel-density data term (the reference's colour model). Restates
 * K_InitGaussian (DataTermKernels.cu:19-31) for the kernel values, and
 * K_SelectAndGauss + K_CalcTerms (DataTermKernels.cu:61-83,116-150) as a
 * per-pixel sum: the volume value pdf[c] after all splats is the sum over
 * seeds s of g[(int)|c - s|] for |c - s| < 101, here summed in double in seed
 * order. floor(|c - s|) is taken with an integer square root. */
static int isqrt_exact(int s) {
    int k = 0;
    while ((k + 1) * (k + 1) <= s) ++k;
    return k;
}

static int32_t trunc_sat(double v) {
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int32_t)0x80000000u;
    return (int32_t)v;
}

void oracle_kde_gauss(int sigma, float *g101) {
    for (int k = 0; k < 101; ++k) {
        const float t = (float)k / (float)sigma;
        const double arg = -0.5 * (double)t * (double)t;
        g101[k] = (float)(1.0 / ((double)(float)sigma * 2.507) * (double)expf((float)arg));
    }
}

int oracle_dataterm_kde(int width, int height, const uint8_t *rgb, int nseeds, const uint8_t *seed_rgb, int sigma,
                        int32_t *pos, int32_t *neg, double *tpos_f, double *tneg_f, double *p_out) {
    if (width <= 0 || height <= 0 || !rgb || nseeds < 0 || (nseeds && !seed_rgb) || sigma < 1) return -1;
    float g[101];
    oracle_kde_gauss(sigma, g);
    int root[101 * 101];
    for (int s = 0; s < 101 * 101; ++s) root[s] = isqrt_exact(s);
    const size_t n = (size_t)width * height;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t *a = rgb + 3 * i;
        double p = 0.0;
        for (int j = 0; j < nseeds; ++j) {
            const uint8_t *b = seed_rgb + 3 * (size_t)j;
            int s = 0;
            for (int c = 0; c < 3; ++c) s += (a[c] - b[c]) * (a[c] - b[c]);
            if (s < 101 * 101) p += (double)g[root[s]];
        }
        const float pf = (float)p;
        if (p_out) p_out[i] = p;
        if (tpos_f) tpos_f[i] = 300000.0 * p;
        if (tneg_f) tneg_f[i] = 300000.0 * (1.0 - p);
        if (pos) pos[i] = trunc_sat(300000.0 * (double)pf);
        if (neg) neg[i] = trunc_sat(300000.0 * (double)(1.0f - pf));
    }
    return 0;
}

/* Optimality certificate of a returned residual state, independent of how it
 * was computed (for sizes where running oracle_maxflow_grid is too slow):
 *  1. residuals are >= 0, zero on every arc into or out of the border, and
 *     r(u,v) + r(v,u) == c(u,v) + c(v,u) for every interior pair;
 *  2. conservation: excess(v) == pos(v) - neg(v) - sum_u (c(v,u) - r(v,u));
 *  3. labels == interior nodes with no residual path to a node of excess < 0
 *     (BFS backwards from those nodes), border labels 0;
 *  4. the returned flow == sum_interior pos - sum_interior max(excess, 0) ==
 *     the cut value of the labels.
 * (1, 2) make the state a preflow of the graph in the reparametrised form
 * (t-links pre-saturated); (3) says no excess node can reach a deficit node,
 * so the preflow's value is a maximum flow and the cut is minimum; the label
 * set is the unique "cannot reach the sink" set. Returns 0 if every check
 * holds, else the number of the first failing check. caps: 4 per pixel
 * (left, right, up, down), image layout. */
int oracle_certify_grid(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                        const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                        const int32_t *caps, const int32_t *excess, const int32_t *label, int64_t flow) {
    const int32_t *planes[8] = {up ? left : NULL, right, up, down, NULL, NULL, NULL, NULL};
    const size_t n = (size_t)width * height;
    if (width < 3 || height < 3) {
        for (size_t i = 0; i < n; ++i)
            if (label[i]) return 3;
        return flow == 0 ? 0 : 4;
    }
    int64_t sp = 0, se = 0;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            const int in = is_interior(x, y, width, height);
            int64_t bal = 0;
            for (int dir = 0; dir < 4; ++dir) {
                const int64_t r = caps[4 * i + dir];
                const int xx = x + DX[dir], yy = y + DY[dir];
                const int open = in && is_interior(xx, yy, width, height);
                if (r < 0 || (!open && r != 0)) return 1;
                if (!open) continue;
                const size_t j = (size_t)yy * width + xx;
                const int64_t c = arc_cap(x, y, width, height, dir, planes, penalty);
                const int64_t cb = arc_cap(xx, yy, width, height, OPP[dir], planes, penalty);
                if (r + caps[4 * j + OPP[dir]] != c + cb) return 1;
                bal += c - r; /* net flow out of i along this pair */
            }
            if (!in) continue;
            if ((int64_t)excess[i] != (int64_t)pos[i] - neg[i] - bal) return 2;
            sp += pos[i];
            se += excess[i] > 0 ? excess[i] : 0;
        }
    /* 3: reverse BFS from deficit nodes over residual arcs u -> v */
    uint8_t *reach = (uint8_t *)calloc(n, 1);
    int32_t *q = (int32_t *)malloc(sizeof(int32_t) * n);
    if (!reach || !q) {
        free(reach), free(q);
        return -2;
    }
    size_t head = 0, tail = 0;
    for (int y = 1; y < height - 1; ++y)
        for (int x = 1; x < width - 1; ++x) {
            const size_t i = (size_t)y * width + x;
            if (excess[i] < 0) reach[i] = 1, q[tail++] = (int32_t)i;
        }
    while (head < tail) {
        const int32_t v = q[head++];
        const int x = v % width, y = v / width;
        for (int dir = 0; dir < 4; ++dir) {
            const int ux = x + DX[dir], uy = y + DY[dir];
            if (!is_interior(ux, uy, width, height)) continue;
            const size_t u = (size_t)uy * width + ux;
            if (reach[u] || caps[4 * u + OPP[dir]] <= 0) continue;
            reach[u] = 1;
            q[tail++] = (int32_t)u;
        }
    }
    int bad = 0;
    for (int y = 0; y < height && !bad; ++y)
        for (int x = 0; x < width; ++x) {
            const size_t i = (size_t)y * width + x;
            const int want = is_interior(x, y, width, height) && !reach[i];
            if (label[i] != want) {
                bad = 1;
                break;
            }
        }
    free(reach), free(q);
    if (bad) return 3;
    if (flow != sp - se) return 4;
    if (cut_core(width, height, pos, neg, planes, 4, penalty, label) != flow) return 4;
    return 0;
}
