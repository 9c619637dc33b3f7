/*
 * gcfg.h -- C ABI of the MI355X-native binary-MRF min-cut library (libgcfg.so).
 *
 * Plain pointers and sizes only; every call returns an int status
 * (GCFG_OK == 0) and gcfg_last_error() gives the message of the last failure
 * on the calling thread. "d_" pointers are device (HBM) pointers allocated by
 * the caller with hipMalloc on the current device; "h_" pointers are host
 * pointers. Directional n-link arrays follow GraphCut.cu:271-289 and
 * GraphCutKernels.cu:186-189: up[x + y*W] is the capacity from pixel (x,y) to
 * (x,y-1), down to (x,y+1), left to (x-1,y), right to (x+1,y). Only interior
 * pixels 0<x<W-1, 0<y<H-1 are graph nodes (GraphCutKernels.cu:122-123,
 * :143-146). Labels: 1 <=> interior node that cannot reach the sink in the
 * final residual graph (SPREAD_ZEROS, GraphCut.h:37), border pixels 0.
 * Input range (checked on the device; a solve of inputs outside it returns
 * GCFG_ERR_INVALID): |pos - neg| <= 2^29 per pixel (t-links of either sign
 * are accepted, as the reference takes them), n-links in [0, 2^28) for the
 * 4-neighbourhood and [0, 2^27) for the 8-neighbourhood. Within it every
 * node's int32 excess and every residual capacity stay in range for any
 * sequence of pushes.
 * Calls on one handle must not overlap; distinct handles are independent.
 */
#ifndef GCFG_H
#define GCFG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GCFG_OK = 0,
    GCFG_ERR_INVALID = -1,       /* bad argument, shape or capacity */
    GCFG_ERR_HIP = -2,           /* HIP runtime error */
    GCFG_ERR_NOMEM = -3,         /* device allocation failed */
    GCFG_ERR_STATE = -4,         /* call out of order (e.g. optimize before build) */
    GCFG_ERR_NOT_CONVERGED = -5, /* step cap reached */
    GCFG_ERR_BUDGET = -6         /* request larger than the serving HBM budget */
};

enum { GCFG_SEED_NONE = 0, GCFG_SEED_FG = 1, GCFG_SEED_BG = 2 };
/* Colour models of the data term: 16-bin seed histograms with -log-likelihood
 * t-links (default), or the reference's Gaussian kernel density over selected
 * colours with pos = (int)(300000 p), neg = (int)(300000 (1 - p))
 * (DataTermKernels.cu:61-83,116-150). */
enum { GCFG_DT_MODEL_HISTOGRAM = 0, GCFG_DT_MODEL_KDE = 1 };

typedef struct gcfg_options {
    int32_t sweeps_per_gr;   /* colour-pass pairs of tile push-relabel between global relabels */
    int32_t inner_iters;     /* max push/relabel iterations per tile pass (default 8; twice as many, at most 64, in exact rounds) */
    int32_t steps_per_poll;  /* scheduler steps launched between two host polls */
    int32_t use_graph;       /* capture a chunk of steps in a hipGraph and replay it */
    int32_t max_steps;       /* 0 = unbounded; else GCFG_ERR_NOT_CONVERGED past this */
    int32_t verbose;         /* 1: print the reference's "pen"/energy lines */
    int32_t profile;         /* 1: time every push/relabel and global-relabel launch with HIP events;
                                2: also count inner iterations (pr_iters, gr_iters; adds atomics) */
    int32_t kernels;         /* 1: wavefront-per-tile kernels (default), 0: workgroup-per-tile LDS kernels */
    int32_t gr_rounds;       /* max tile-local relaxation rounds per global-relabel sweep (then re-queued;
                                default 48) */
    int32_t gr_block_w;      /* global-relabel block in tiles (one workgroup, BFS exchanged through LDS): */
    int32_t gr_block_h;      /*   1x1, 1x2, 1x4, 1x8, 1x16, 2x2, 2x4, 2x8 or 4x4; 0 = default */
    int32_t gr_cap;          /* >0 (default 4): the first global relabel of an instance (the one the
                                build runs) stops BFS fronts at this distance (nodes beyond are
                                frozen until the next relabel), the next three at gr_cap_late, later
                                ones are exact; exact rounds after the first recompute only the
                                blocks near the pushes since the last round, and an instance ends
                                with a round over every block (reachability after capped rounds,
                                else exact); 0: every relabel exact over every block */
    int32_t pr_waves_per_cu; /* push/relabel launch size in wavefronts per compute unit (0 = as many as stay resident) */
    int32_t gr_waves_per_cu; /* global-relabel launch size in wavefronts per compute unit (0 = 48) */
    int32_t lazy_caps;       /* residual capacities of the wavefront kernels: 0 = chosen per build (lazy for
                                large batches with few steps), 1 = written by the build, 2 = lazy (written
                                before a tile's first push pass) */
    int32_t gr_cap_late;     /* cap of the second to fourth global relabels when gr_cap > 0 (default 15;
                                0 or anything below gr_cap: gr_cap) */
} gcfg_options;

typedef struct gcfg_stats {
    int64_t flow;        /* exact max-flow value over the interior */
    int64_t pr_tiles;    /* tile push-relabel passes executed */
    int64_t gr_tiles;    /* tile BFS sweeps executed (global relabel) */
    int64_t pr_iters;    /* in-LDS push/relabel iterations summed over tiles (profile 2) */
    int32_t steps;       /* scheduler steps until done */
    int32_t gr_rounds;   /* global relabels started */
    double ms;           /* wall time of the call */
    /* filled when gcfg_options.profile is set (HIP events around each launch) */
    double pr_kernel_ms, gr_kernel_ms, sched_kernel_ms;
    int64_t launches;    /* steps launched (each = sched + pr + gr) */
    int64_t gr_iters;    /* tile-local relaxation rounds summed over global-relabel tiles (profile 2) */
    int64_t gr_init_tiles; /* of gr_tiles, the initial sweeps of a global relabel (the rest relax) */
    int64_t gr_exact_tiles; /* of gr_tiles, those of exact (uncapped) rounds, which also write labels */
    int64_t fin_items;     /* FIN items: blocks left unreachable (wavefront kernels) or every tile */
    double build_kernel_ms; /* graph build kernel (GC_SetGraph's work), when profile is set */
    /* gr_* and fin_items count relabel items: blocks of gr_block_w x gr_block_h tiles */
} gcfg_stats;

void gcfg_default_options(gcfg_options *opt);
const char *gcfg_last_error(void);
/* Name, compute-unit count and HBM size of the current device. */
int gcfg_device_info(char *name, size_t name_len, int32_t *cu_count, size_t *hbm_bytes);

/* ---- Plain drop-in graph: the GC_* functions of GraphCut.cu ---- */
typedef struct gcfg_graph gcfg_graph;
/* GC_Init allocation part, GraphCut.cu:187-248. */
int gcfg_graph_create(int32_t width, int32_t height, gcfg_graph **out);
/* GC_Init with NEIGHBORHOOD == 8 (GraphCut.cu:187-194,223-232; GraphCut.h:22,65-70):
 * an 8-connected graph. Its n-links come from gcfg_graph_set_edges8 or
 * gcfg_graph_set_penalty (all eight directions); every other call is as for
 * the 4-neighbourhood. Solved by the workgroup-per-tile kernels
 * (gcfg_options.kernels is ignored). */
int gcfg_graph_create_n8(int32_t width, int32_t height, gcfg_graph **out);
/* 4 or 8. */
int gcfg_graph_neighborhood(gcfg_graph *g, int32_t *nbhd);
/* GC_SetDataterms, GraphCut.cu:265-269 (borrowed device pointers, W*H each). */
int gcfg_graph_set_dataterms(gcfg_graph *g, const int32_t *d_pos, const int32_t *d_neg);
/* GC_SetEdges, GraphCut.cu:271-289 (borrowed device pointers, W*H each). */
int gcfg_graph_set_edges(gcfg_graph *g, const int32_t *d_up, const int32_t *d_down, const int32_t *d_left,
                         const int32_t *d_right);
/* GC_SetEdges with NEIGHBORHOOD == 8, GraphCut.cu:271-289: the four side planes
 * and the diagonals, upleft[x+y*W] = capacity from (x,y) to (x-1,y-1),
 * upright to (x+1,y-1), downleft to (x-1,y+1), downright to (x+1,y+1)
 * (borrowed device pointers, W*H each; 8-neighbourhood graphs only). */
int gcfg_graph_set_edges8(gcfg_graph *g, const int32_t *d_up, const int32_t *d_down, const int32_t *d_left,
                          const int32_t *d_right, const int32_t *d_upleft, const int32_t *d_upright,
                          const int32_t *d_downleft, const int32_t *d_downright);
/* GC_SetPenalty, GraphCut.cu:296-299 (uniform n-link capacity). */
int gcfg_graph_set_penalty(gcfg_graph *g, int32_t penalty);
/* Order the handle's work after everything queued on a caller HIP stream
 * (hipStream_t passed as void*) at each build/solve, and make that stream
 * wait for the handle's work (device labels included) when the call returns.
 * NULL restores the default (calls are synchronous to the host only). The
 * reference runs everything on the legacy default stream (GraphCut.cu). */
int gcfg_graph_set_stream(gcfg_graph *g, void *hip_stream);
int gcfg_graph_set_options(gcfg_graph *g, const gcfg_options *opt);
/* GC_SetGraph, GraphCut.cu:301-319: builds the residual graph on the device. */
int gcfg_graph_build(gcfg_graph *g);
/* GC_Optimize, GraphCut.cu:385-534: solves and writes W*H labels to host memory. */
int gcfg_graph_optimize(gcfg_graph *g, int32_t *h_label, gcfg_stats *stats);
/* Same solve with the W*H labels left in caller-owned device memory. */
int gcfg_graph_optimize_device(gcfg_graph *g, int32_t *d_label, gcfg_stats *stats);
/* Residual state after optimize, image layout: h_caps[4*(x+y*W)+k] is the
 * residual capacity from (x,y) towards left/right/up/down (k = 0..3),
 * h_excess[x+y*W] the excess (negative = remaining sink capacity). */
int gcfg_graph_read_state(gcfg_graph *g, int32_t *h_caps, int32_t *h_excess);
/* Diagonal residual capacities of an 8-neighbourhood graph after optimize:
 * h_caps8[4*(x+y*W)+k], k = 0..3 towards up-left, up-right, down-left,
 * down-right. Verification output with no reference counterpart. */
int gcfg_graph_read_diag_state(gcfg_graph *g, int32_t *h_caps8);
/* Distance labels (W*H, image layout) after a solve: the exact BFS distance
 * of every interior node to a node with excess < 0 along residual arcs in the
 * final residual graph (gr_cap 0, the workgroup kernels, or an instance that
 * outlasted its capped rounds); 0x3fffffff where no such path exists. An
 * instance that finished within capped rounds ends with a reachability
 * round instead: 0 for the nodes with such a path, 0x3fffffff for the others.
 * No counterpart in the reference (its heights are approximate); exposed for
 * verification. */
int gcfg_graph_read_distances(gcfg_graph *g, int32_t *h_dist);
/* GC_End, GraphCut.cu:542-568 (borrowed inputs are not freed). */
int gcfg_graph_destroy(gcfg_graph *g);

/* ---- Colour data term: the DataTerm.h functions ---- */
typedef struct gcfg_dataterm gcfg_dataterm;
/* init_data_term, DataTerm.cpp:28-65. */
int gcfg_dt_create(int32_t width, int32_t height, gcfg_dataterm **out);
/* free_data_term, DataTerm.cpp:67-78. */
int gcfg_dt_destroy(gcfg_dataterm *dt);
/* setImage, DataTerm.cpp:80-86 (W*H*3 interleaved u8, host). */
int gcfg_dt_set_image(gcfg_dataterm *dt, const uint8_t *h_rgb);
/* Same with a caller-owned device image. */
int gcfg_dt_set_image_device(gcfg_dataterm *dt, const uint8_t *d_rgb);
/* setSelection, DataTerm.cpp:139-153: per-pixel seed classes (GCFG_SEED_*), host. */
int gcfg_dt_set_seeds(gcfg_dataterm *dt, const uint8_t *h_seeds);
/* selectOnMap, DataTerm.cpp:105-127: mark a w x h rectangle of seeds of class cls. */
int gcfg_dt_add_seed_rect(gcfg_dataterm *dt, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t cls);
/* reset, DataTerm.cpp:88-95: clear all seeds. */
int gcfg_dt_reset(gcfg_dataterm *dt);
/* n-link weight lambda (default 50). */
int gcfg_dt_set_lambda(gcfg_dataterm *dt, int32_t lambda);
/* getDataTermsMixture, DataTerm.cpp:129-137: histograms, beta, t-links, n-links. */
int gcfg_dt_compute(gcfg_dataterm *dt);
/* As gcfg_graph_set_stream, for gcfg_dt_compute. */
int gcfg_dt_set_stream(gcfg_dataterm *dt, void *hip_stream);
/* getPos / getNeg, DataTerm.cpp:12-26, plus the four n-link planes (device, owned by dt). */
int gcfg_dt_get_links(gcfg_dataterm *dt, int32_t **d_pos, int32_t **d_neg, int32_t **d_up, int32_t **d_down,
                      int32_t **d_left, int32_t **d_right);
/* Seed histograms (class-major: fg then bg; channel; 16 bins) and seed counts. */
int gcfg_dt_get_histograms(gcfg_dataterm *dt, int32_t *h_hist96, int32_t *h_counts2);
/* Exact sum of squared 4-neighbour colour differences, pair count, beta. */
int gcfg_dt_get_beta(gcfg_dataterm *dt, int64_t *sum, int64_t *pairs, double *beta);
/* Pre-quantisation values computed on the device in f64 (W*H each, host out).
 * Kernel-density model: h_tpos = 300000 p, h_tneg = 300000 (1 - p). */
int gcfg_dt_unquantized(gcfg_dataterm *dt, double *h_tpos, double *h_tneg, double *h_right, double *h_down);
/* Select the colour model (GCFG_DT_MODEL_*) used by gcfg_dt_compute. No
 * counterpart in the reference, whose model is the kernel density only. */
int gcfg_dt_set_model(gcfg_dataterm *dt, int32_t model);
/* setSigmaGaussian, DataTerm.cpp:97-103 / K_InitGaussian, DataTermKernels.cu:19-31:
 * kernel width (1..1000, default 5) of the density; applies to later computes. */
int gcfg_dt_set_sigma(gcfg_dataterm *dt, int32_t sigma);
/* selectOnMap, DataTerm.cpp:105-127 / K_SelectAndGauss, DataTermKernels.cu:116-150:
 * adds the colours of the 5 x 5 pixels (x..x+4, y..y+4) of the current image
 * to the kernel-density seed list (pixels outside the image are skipped; a
 * pixel selected twice counts twice). Later image changes do not change the
 * captured colours. At most 65536 seed colours; gcfg_dt_reset clears them. */
int gcfg_dt_kde_select(gcfg_dataterm *dt, int32_t x, int32_t y);
/* Number of seed colours in the kernel-density list. */
int gcfg_dt_kde_count(gcfg_dataterm *dt, int32_t *count);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gcfg_dt_compute's histogram data term (lambda, 16-bin
 * seed histograms; same outputs) of one image of at most
 * GCFG_DT_NOALLOC_MAX_PIXELS pixels, from a caller-owned device image
 * (W*H*3, 4-byte aligned) and seed map (W*H, 4-byte aligned, or NULL) into
 * six caller-owned W*H int32 planes, with every scratch buffer in a fixed
 * device array of the library (161 KiB per GPU, held until the process ends).
 * It allocates nothing, issues no memset and does not synchronise: the work
 * is queued on hip_stream (NULL: the default stream) and the outputs are
 * ready when the stream reaches it. Calls that overlap on one GPU share the
 * scratch and race; larger images fail with GCFG_ERR_INVALID. */
#define GCFG_DT_NOALLOC_MAX_PIXELS (1 << 26)
int gcfg_dt_compute_noalloc(int32_t width, int32_t height, const uint8_t *d_rgb, const uint8_t *d_seeds,
                            int32_t lambda, int32_t *d_pos, int32_t *d_neg, int32_t *d_up, int32_t *d_down,
                            int32_t *d_left, int32_t *d_right, void *hip_stream);

/* Batched data term: getDataTermsMixture (DataTerm.cpp:129-137) of many
 * images at once, histogram model, with one launch per pass for the whole
 * batch (no counterpart in the reference, which computes one image per
 * call). Each image: caller-owned device RGB (W*H*3 interleaved, 4-byte
 * aligned), its seed map (W*H classes GCFG_SEED_*, 4-byte aligned; NULL: no
 * seeds) and six output planes (W*H int32 each; 16-byte aligned planes are
 * written with vector stores). Outputs equal gcfg_dt_compute's per image. */
typedef struct gcfg_dt_image {
    int32_t width, height;
    const uint8_t *d_rgb;
    const uint8_t *d_seeds;
    int32_t *d_pos, *d_neg, *d_up, *d_down, *d_left, *d_right;
} gcfg_dt_image;
typedef struct gcfg_dt_batch gcfg_dt_batch;
int gcfg_dt_batch_create(int32_t count, const gcfg_dt_image *images, int32_t lambda, gcfg_dt_batch **out);
/* As gcfg_graph_set_stream, for gcfg_dt_batch_compute. */
int gcfg_dt_batch_set_stream(gcfg_dt_batch *b, void *hip_stream);
/* Computes every image's data term; returns when the outputs are written. */
int gcfg_dt_batch_compute(gcfg_dt_batch *b);
int gcfg_dt_batch_destroy(gcfg_dt_batch *b);

/* ---- Batched solve: many instances of uneven size in one engine ---- */
typedef struct gcfg_instance {
    int32_t width, height;
    const int32_t *d_pos, *d_neg;                   /* W*H each */
    const int32_t *d_up, *d_down, *d_left, *d_right; /* W*H each, or all NULL for a uniform penalty */
    int32_t penalty;
} gcfg_instance;
typedef struct gcfg_batch gcfg_batch;
int gcfg_batch_create(int32_t count, const gcfg_instance *inst, const gcfg_options *opt, gcfg_batch **out);
/* As gcfg_graph_set_stream, for gcfg_batch_solve / gcfg_batch_solve_device. */
int gcfg_batch_set_stream(gcfg_batch *b, void *hip_stream);
/* Cold solve of every instance; labels to host (h_labels[i]: W_i*H_i) and flows. */
int gcfg_batch_solve(gcfg_batch *b, int32_t **h_labels, int64_t *flows, gcfg_stats *stats);
/* Cold solve with labels written to caller-owned device buffers (d_labels[i]: W_i*H_i). */
int gcfg_batch_solve_device(gcfg_batch *b, int32_t **d_labels, int64_t *flows, gcfg_stats *stats);
/* Cold solve with the labels on the host packed one bit per pixel: h_bits[i]
 * receives a pointer to instance i's map in pinned host memory owned by the
 * batch, valid until the next solve or destroy. Row y of instance i is the
 * ceil(W_i/64) words h_bits[i][y * ceil(W_i/64) + k], bit x of word k being
 * the label of pixel (64 k + x, y); bits past W_i are 0. The GC_Optimize
 * result (GraphCut.cu:528-529: int32 labels copied to the host) in 1/32 of
 * the bytes, written to the host as each instance finishes (by the device,
 * while the other instances are still being solved), so no transfer is left
 * after the solve. */
int gcfg_batch_solve_bits(gcfg_batch *b, const uint64_t **h_bits, int64_t *flows, gcfg_stats *stats);
int gcfg_batch_destroy(gcfg_batch *b);

/* ---- File formats of the reference's drivers (host only) ----
 * Middlebury 2-label MRF text files: gcfg_mrf_info reads the header
 * (solveMRF.cu:16); gcfg_mrf_load fills caller arrays of W*H int32 with the
 * data terms and the cue planes mapped as the reference does
 * (solveMRF.cu:10-65, 87-90: up = down = vertical cue, left = right =
 * horizontal cue, last row of both cues zeroed); gcfg_mrf_write writes one,
 * repeated `mult` times horizontally (multiplyMRF.cpp:64-92). */
int gcfg_mrf_info(const char *path, int32_t *width, int32_t *height, int32_t *n_labels);
int gcfg_mrf_load(const char *path, int32_t width, int32_t height, int32_t *pos, int32_t *neg, int32_t *up,
                  int32_t *down, int32_t *left, int32_t *right);
int gcfg_mrf_write(const char *path, int32_t width, int32_t height, int32_t mult, const int32_t *pos,
                   const int32_t *neg, const int32_t *hcue, const int32_t *vcue);
/* Data-term files (rihanna.dat / face.dat of Application.cu:15-24 and
 * multiplyMRF.cpp:130-145): "height width", then all W*H neg values, then all
 * W*H pos values, row-major, whitespace separated. */
int gcfg_dat_info(const char *path, int32_t *width, int32_t *height);
int gcfg_dat_load(const char *path, int32_t width, int32_t height, int32_t *pos, int32_t *neg);
int gcfg_dat_write(const char *path, int32_t width, int32_t height, const int32_t *pos, const int32_t *neg);
/* writeRihannaMRFData (multiplyMRF.cpp:94-122): a Middlebury 2-label MRF file
 * of out_w x out_h pixels (the reference: 3600 x 3600) whose data terms repeat
 * the width x height source (pixel (x, y) takes source pixel
 * (x mod width, y mod height); the reference indexes rows without the wrap,
 * which agrees wherever it reads inside the source) and whose horizontal and
 * vertical cues all equal `penalty` (the reference: 600). */
int gcfg_mrf_write_tiled(const char *path, int32_t out_w, int32_t out_h, int32_t width, int32_t height,
                         const int32_t *pos, const int32_t *neg, int32_t penalty);
/* PPM input images (P3 or P6, maxval 255; Application.cu:55-58) as
 * interleaved RGB u8. */
int gcfg_ppm_info(const char *path, int32_t *width, int32_t *height);
int gcfg_ppm_read(const char *path, int32_t width, int32_t height, uint8_t *rgb);
/* Label images: text P3 with white (rgb NULL, solveMRF.cu:113-127) or the
 * input colour (Application.cu:61-75) where label is 1, black elsewhere; and
 * binary P5 with 255 where label is 1. */
int gcfg_write_label_ppm(const char *path, int32_t width, int32_t height, const int32_t *label, const uint8_t *rgb);
int gcfg_write_label_pgm(const char *path, int32_t width, int32_t height, const int32_t *label);

/* ---- Autotuning of the solver options ----
 * Coordinate search over inner_iters, sweeps_per_gr, gr_rounds and
 * steps_per_poll with cold solves of the handle's own inputs (the fastest of
 * `trials` solves per candidate); the winner is applied to the handle and
 * returned. Results are bit-identical for every candidate. */
int gcfg_graph_autotune(gcfg_graph *g, int32_t trials, gcfg_options *best, double *best_ms);
int gcfg_batch_autotune(gcfg_batch *b, int32_t trials, gcfg_options *best, double *best_ms);
/* Persisted choices: one "key value..." line per workload class in a text file. */
int gcfg_tune_store(const char *path, const char *key, const gcfg_options *opt);
/* GCFG_ERR_STATE if the key is absent. */
int gcfg_tune_lookup(const char *path, const char *key, gcfg_options *opt);

/* ---- Serving: continuous batching under an HBM budget, warm-started edits ---- */
typedef struct gcfg_server gcfg_server;
/* Label layouts of served requests: W*H int32 (the GC_Optimize layout), or
 * packed one bit per pixel as gcfg_batch_solve_bits writes them (H rows of
 * ceil(W/64) uint64 words; h_label then points to that many words). */
enum { GCFG_LABELS_INT32 = 0, GCFG_LABELS_BITS = 1 };
typedef struct gcfg_result {
    int64_t request_id;
    int32_t session;
    int32_t status;      /* GCFG_OK or an error code */
    int64_t flow;
    double latency_ms;   /* submit -> labels in host memory */
    double solve_ms;     /* admission -> labels in host memory */
    double done_ms;      /* completion time on the gcfg_now_ms() clock */
    /* Requests submitted with h_label NULL: the labels in the server's pinned
     * host memory (label_bytes bytes, the request's label format), valid until
     * the next gcfg_server_poll on this server; NULL otherwise. */
    const void *labels;
    size_t label_bytes;
} gcfg_result;
/* Monotonic clock (ms) shared with gcfg_result.done_ms. */
double gcfg_now_ms(void);
int gcfg_server_create(size_t hbm_budget_bytes, const gcfg_options *opt, gcfg_server **out);
/* Label layout (GCFG_LABELS_*) of the requests submitted after this call
 * (default GCFG_LABELS_INT32). No counterpart in the reference. */
int gcfg_server_set_label_format(gcfg_server *s, int32_t format);
/* New session: image + seed map (host), data term and solve on the device.
 * h_label (host, in the server's label format) receives the labels; with
 * h_label NULL the result lends them instead (gcfg_result.labels: no copy
 * into caller memory on the request path). h_rgb, h_seeds and h_label must
 * stay valid, and the inputs unchanged, until the request's result is
 * returned by gcfg_server_poll (the image is copied to
 * the device asynchronously; from pinned memory the copy does not hold up
 * the calling thread). */
int gcfg_server_submit_image(gcfg_server *s, int32_t width, int32_t height, const uint8_t *h_rgb,
                             const uint8_t *h_seeds, int32_t *h_label, int64_t *request_id, int32_t *session);
/* Edit of a live session: n strokes of 5 ints (x0, y0, w, h, cls); a
 * foreground stroke adds delta to pos, a background stroke adds delta to neg,
 * on the stroke pixels. Warm-starts from the device-resident residual graph.
 * A session takes edits while the sum of delta x n_strokes over them stays at
 * most 2^29; the request past that fails with GCFG_ERR_INVALID (the int32
 * excess bound below). */
int gcfg_server_submit_edit(gcfg_server *s, int32_t session, int32_t n_strokes, const int32_t *strokes,
                            int32_t delta, int32_t *h_label, int64_t *request_id);
/* Next video frame of a live session (the loop of ComplexApplication.cu:202-207,
 * which re-runs setImage / getDataTermsMixture / GC_SetGraph / GC_Optimize per
 * frame): a W x H RGB image of the session's size (host, valid and unchanged
 * until the request's result is returned). The data term is recomputed on the device with the
 * colour model of the session's first image, and the device-resident residual
 * graph is reparametrised onto the new capacities (flow beyond a new n-link
 * capacity returns to its tail node) instead of rebuilt, then solved warm.
 * Strokes of earlier edits stay applied. Labels and flow are those of a cold
 * solve of the same capacities. */
int gcfg_server_submit_frame(gcfg_server *s, int32_t session, int32_t width, int32_t height, const uint8_t *h_rgb,
                             int32_t *h_label, int64_t *request_id);
/* Advance the device and collect finished requests. wait == 0: one pass that
 * never waits for the device (it admits queued requests, launches the next
 * chunk of solver steps and returns what has finished); wait != 0: until at
 * least one result is ready or nothing is left. A request's labels are in
 * its h_label when its result is returned. */
int gcfg_server_poll(gcfg_server *s, int32_t wait, int32_t max_results, gcfg_result *results, int32_t *n_results);
/* Drop a session and return its memory to the budget. */
int gcfg_server_close_session(gcfg_server *s, int32_t session);
/* Requests admitted and waiting, and HBM bytes in use. */
int gcfg_server_load(gcfg_server *s, int32_t *in_flight, int32_t *queued, size_t *bytes_in_use);
int gcfg_server_destroy(gcfg_server *s);

/* ---- Seeded synthetic inputs (host-only, no GPU work) ---- */
int gcfg_gen_ellipse_image(int width, int height, uint64_t seed, uint8_t *rgb);
int gcfg_gen_seed_map(int width, int height, uint8_t *seeds);
int gcfg_gen_stress_tlinks(int width, int height, uint64_t seed, int32_t *pos, int32_t *neg);
int gcfg_gen_perlin_image(int width, int height, uint64_t seed, uint8_t *rgb);
int gcfg_gen_batch_shapes(int count, uint64_t seed, double area_min, double area_max, int32_t *widths,
                          int32_t *heights);
uint64_t gcfg_gen_mix_seed(uint64_t seed, uint64_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GCFG_H */
