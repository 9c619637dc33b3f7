/*
 * GraphCut.h -- drop-in C++ host API of the reference solver, backed by
 * libgcfg.so (MI355X HIP kernels through the C ABI of gcfg.h).
 *
 * Keeps the names, argument order, default arguments and pass-by-value /
 * pass-by-pointer conventions of the reference (GraphCut.h:60-128 structs,
 * GraphCut.cu:187-568 functions) with SPREAD_ZEROS. NEIGHBORHOOD is 4 unless
 * the includer defines it as 8 (the reference edits GraphCut.h:22); with 8,
 * GC_Init and GC_SetEdges take the four diagonal planes as well
 * (GraphCut.cu:190-193,273-276) and the graph is 8-connected. The library
 * exports both arities.
 * Inputs are caller-owned device pointers (int32, W*H, row-major x + y*W),
 * borrowed until GC_Optimize returns; labels go to a caller-owned host
 * int[W*H]. As in the reference, a failing call prints a message and aborts.
 * Differences: the compile-time tuning macros of GraphCut.h:39-171 are
 * runtime options (GC_SetOptions); GC_GetFlow returns the exact int64
 * max-flow of the last GC_Optimize (the reference only prints an energy).
 */
#ifndef GRAPHCUT_H_
#define GRAPHCUT_H_

#include <cstddef>

#include "gcfg.h"

#ifndef NEIGHBORHOOD
#define NEIGHBORHOOD 4
#endif
#if NEIGHBORHOOD != 4 && NEIGHBORHOOD != 8
#error "NEIGHBORHOOD must be 4 or 8"
#endif
#define SPREAD_ZEROS

/* The reference's wrapper structs (GraphCut.h:60-128) with their field names,
 * so that code reading them compiles. The geometry fields describe this
 * library's padded tile grid (tiles of 64 x 16 nodes, the counterpart of the
 * reference's 32 x 8 thread blocks): width_ex, height_ex and size_ex are the
 * padded extents, block_x x block_y the tile grid and block_count the tiles.
 * The residual state lives in the library's tiled arena, not in W*H planes:
 * the NodeWrapper pointers and KernelWrapper::active are NULL (the state is
 * read with gcfg_graph_read_state / gcfg_graph_read_distances). The diagonal
 * fields, under NEIGHBORHOOD == 8 only in the reference, are present in both
 * builds: one library serves both arities, so the layout is the same. */
struct NodeWrapper {
    int *edge_l;
    int *edge_r;
    int *edge_u;
    int *edge_d;
    int *edge_ul;
    int *edge_dr;
    int *edge_ur;
    int *edge_dl;
    int *height;
    int *excess;
    int *status;
    int *comp_h;
    int *comp_n;
    int *energy_sum;
};

struct GraphWrapper {
    int width;     /* graph width */
    int height;    /* graph height */
    int size;      /* width * height */
    int width_ex;  /* width padded to whole tiles */
    int height_ex; /* height padded to whole tiles */
    int size_ex;   /* width_ex * height_ex */
    NodeWrapper n;
};

struct KernelWrapper {
    int block_x; /* tiles across */
    int *active;
    GraphWrapper g;
};

/* Handle plus the borrowed inputs (GraphCut.h:110-128); the last three
 * fields are this library's. */
struct GlobalWrapper {
    int block_y;           /* tiles down */
    int block_count;       /* tiles */
    int *data_positive;    /* source t-links (device) */
    int *data_negative;    /* sink t-links (device) */
    int *up;               /* capacity to (x, y-1) (device) */
    int *down;             /* capacity to (x, y+1) (device) */
    int *left;             /* capacity to (x-1, y) (device) */
    int *right;            /* capacity to (x+1, y) (device) */
    int *upleft;           /* capacity to (x-1, y-1) (device; 8-neighbourhood) */
    int *upright;          /* capacity to (x+1, y-1) */
    int *downleft;         /* capacity to (x-1, y+1) */
    int *downright;        /* capacity to (x+1, y+1) */
    int penalty;           /* uniform n-link capacity when varying_edges is false */
    bool varying_edges;    /* whether per-direction capacities were given */
    KernelWrapper k;       /* geometry (k.g.width, k.g.height, ...) */
    int neighborhood;      /* 4 or 8, fixed by GC_Init */
    gcfg_graph *graph;     /* device state */
    long long *last_flow;  /* exact max-flow of the last GC_Optimize (shared by copies) */
};

#if NEIGHBORHOOD == 8
GlobalWrapper GC_Init(int width, int height, int *data_positive = NULL, int *data_negative = NULL, int penalty = 0,
                      int *up = NULL, int *down = NULL, int *left = NULL, int *right = NULL, int *upleft = NULL,
                      int *upright = NULL, int *downleft = NULL, int *downright = NULL);
void GC_SetEdges(GlobalWrapper *gw, int *up = NULL, int *down = NULL, int *left = NULL, int *right = NULL,
                 int *upleft = NULL, int *upright = NULL, int *downleft = NULL, int *downright = NULL);
#else
GlobalWrapper GC_Init(int width, int height, int *data_positive = NULL, int *data_negative = NULL, int penalty = 0,
                      int *up = NULL, int *down = NULL, int *left = NULL, int *right = NULL);
void GC_SetEdges(GlobalWrapper *gw, int *up = NULL, int *down = NULL, int *left = NULL, int *right = NULL);
#endif
void GC_SetDataterms(GlobalWrapper *gw, int *data_positive, int *data_negative);
void GC_SetPenalty(GlobalWrapper *gw, int p);
void GC_SetGraph(GlobalWrapper gw);
void GC_Optimize(GlobalWrapper gw, int *label);
void GC_End(GlobalWrapper *gw);

/* Additions over the reference API. */
void GC_SetOptions(GlobalWrapper gw, const gcfg_options *opt);
long long GC_GetFlow(GlobalWrapper gw);

#endif /* GRAPHCUT_H_ */
