/*
 * CPU oracle for the binary-MRF min-cut path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker. The product (gc-fg_amd) never links
 * or calls it.
 *
 * Contents:
 *  - oracle_maxflow_grid: exact int64 max-flow / min-cut on the 4-connected
 *    grid graph built with the reference's rules (GraphCutKernels.cu:112-214):
 *    only interior pixels 0<x<W-1, 0<y<H-1 are nodes, every arc towards a
 *    border pixel has capacity 0, t-links are pos (source) and neg (sink).
 *    Labels follow GraphCutKernels.cu:228-351 with SPREAD_ZEROS
 *    (GraphCut.h:37): 1 <=> interior node that cannot reach the sink in the
 *    final residual graph; border pixels are 0.
 *  - oracle_cut_value: the cut value of a labelling (certificate side).
 *  - oracle_dataterm: the colour data term (16-bin per-channel seed
 *    histograms, -log-likelihood t-links, contrast n-links), evaluated per
 *    pixel in double precision, with pre-quantisation values.
 *  - oracle_dataterm_kde: the reference's own Gaussian kernel-density colour
 *    model, as a per-pixel double sum over the selected seed colours.
 *  - oracle_refpr_*: a CPU restatement of the reference's own push-relabel
 *    schedule (GraphCut.cu:385-534 over GraphCutKernels.cu), used as the
 *    timed CPU baseline.
 */
#ifndef GCFG_ORACLE_H
#define GCFG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Directional arrays follow GraphCut.cu:271-289 / GraphCutKernels.cu:186-189:
 * up[x+y*W] is the capacity from (x,y) to (x,y-1), down to (x,y+1),
 * left to (x-1,y), right to (x+1,y). If up is NULL, every n-link is penalty
 * (GraphCutKernels.cu:129-132). Returns 0 on success. */
int oracle_maxflow_grid(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                        const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                        int32_t *label, int64_t *flow);

/* 8-neighbourhood graph (NEIGHBORHOOD 8, GraphCut.h:22,65-70): dirs8 holds
 * eight W*H planes in the order left, right, up, down, up-left (to
 * (x-1,y-1)), up-right (x+1,y-1), down-left (x-1,y+1), down-right (x+1,y+1),
 * or is NULL for a uniform penalty on all eight (GraphCutKernels.cu:129-151).
 * Arcs into the border are zero as in GraphCutKernels.cu:143-156,200-213. */
int oracle_maxflow_grid8(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *const *dirs8,
                         int32_t penalty, int32_t *label, int64_t *flow);
int64_t oracle_cut_value8(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *const *dirs8,
                          int32_t penalty, const int32_t *label);

/* Optimality certificate of a residual state returned by a solver (caps: 4
 * per pixel, left/right/up/down; excess; labels; flow): feasibility,
 * conservation, labels == nodes that cannot reach a deficit node, flow ==
 * sum pos - sum max(excess, 0) == cut of the labels. 0 if it holds, else the
 * number (1-4) of the first failing check. */
int oracle_certify_grid(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                        const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                        const int32_t *caps, const int32_t *excess, const int32_t *label, int64_t flow);

/* oracle_maxflow_grid that also returns its final residual state in the
 * solver's form: caps_out[4*i + k] (left, right, up, down), exc_out[i] =
 * excess (negative: remaining sink capacity). */
int oracle_maxflow_grid_state(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                              const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                              int32_t *label, int64_t *flow, int32_t *caps_out, int32_t *exc_out);

/* Cut value of a 0/1 labelling (1 = source side) on the same graph. */
int64_t oracle_cut_value(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                         const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                         const int32_t *label);

/* Colour data term. rgb: W*H*3 interleaved u8; seeds: W*H, 1 = foreground,
 * 2 = background, anything else unseeded. Outputs (each may be NULL):
 *   hist[2*3*16] (class-major, then channel, then bin; class 0 = foreground),
 *   counts[2], beta_sum (sum of squared colour differences over all
 *   4-neighbour pairs), pairs, beta, pos/neg/up/down/left/right (int32 W*H),
 *   tpos_f/tneg_f (64*-ln p before rounding), right_f/down_f (lambda*exp(..)
 *   before rounding; 0 where there is no neighbour). */
int oracle_dataterm(int width, int height, const uint8_t *rgb, const uint8_t *seeds, int32_t lambda,
                    int32_t *hist, int32_t *counts, int64_t *beta_sum, int64_t *pairs, double *beta,
                    int32_t *pos, int32_t *neg, int32_t *up, int32_t *down, int32_t *left, int32_t *right,
                    double *tpos_f, double *tneg_f, double *right_f, double *down_f);

/* As oracle_dataterm, with the histograms taken from the seed pixels of
 * model_rgb (a video session's first frame) and everything else from rgb. */
int oracle_dataterm_model(int width, int height, const uint8_t *rgb, const uint8_t *model_rgb,
                          const uint8_t *seeds, int32_t lambda, int32_t *hist, int32_t *counts, int64_t *beta_sum,
                          int64_t *pairs, double *beta, int32_t *pos, int32_t *neg, int32_t *up, int32_t *down,
                          int32_t *left, int32_t *right, double *tpos_f, double *tneg_f, double *right_f,
                          double *down_f);

/* Kernel-density colour model of the reference (DataTermKernels.cu:19-31,
 * 61-83, 116-150): g101 receives the 101 kernel values for sigma; the data
 * term sums g[floor(|I - s|)] over the nseeds seed colours (seed_rgb, 3 bytes
 * each, in selection order) with |I - s| < 101 in double; p_out/tpos_f/tneg_f
 * get p, 300000 p and 300000 (1 - p) before truncation; pos/neg are
 * (int)(300000 pf) and (int)(300000 (1 - pf)) with pf = (float)p, saturated
 * to int32. Each output may be NULL. */
void oracle_kde_gauss(int sigma, float *g101);
int oracle_dataterm_kde(int width, int height, const uint8_t *rgb, int nseeds, const uint8_t *seed_rgb, int sigma,
                        int32_t *pos, int32_t *neg, double *tpos_f, double *tneg_f, double *p_out);

/* Reference-schedule push-relabel (CPU restatement). threads <= 0 uses the
 * OpenMP default. Returns the number of outer iterations, or < 0 on error. */
int oracle_refpr_solve(int width, int height, const int32_t *pos, const int32_t *neg, const int32_t *up,
                       const int32_t *down, const int32_t *left, const int32_t *right, int32_t penalty,
                       int threads, int32_t *label, int64_t *flow, int64_t *printed_energy);

#ifdef __cplusplus
}
#endif

#endif
