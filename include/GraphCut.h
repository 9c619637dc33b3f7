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

/* Handle plus the borrowed inputs, mirroring GlobalWrapper (GraphCut.h:110-128). */
struct GlobalWrapper {
    int width;             /* graph width (GraphWrapper::width) */
    int height;            /* graph height (GraphWrapper::height) */
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
