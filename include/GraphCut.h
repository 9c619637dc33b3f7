/*
 * GraphCut.h -- drop-in C++ host API of the reference solver, backed by
 * libgcfg.so (MI355X HIP kernels through the C ABI of gcfg.h).
 *
 * Keeps the names, argument order, default arguments and pass-by-value /
 * pass-by-pointer conventions of the reference (GraphCut.h:60-128 structs,
 * GraphCut.cu:187-568 functions) for NEIGHBORHOOD 4 with SPREAD_ZEROS.
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

#define NEIGHBORHOOD 4
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
    int penalty;           /* uniform n-link capacity when varying_edges is false */
    bool varying_edges;    /* whether per-direction capacities were given */
    gcfg_graph *graph;     /* device state */
    long long *last_flow;  /* exact max-flow of the last GC_Optimize (shared by copies) */
};

GlobalWrapper GC_Init(int width, int height, int *data_positive = NULL, int *data_negative = NULL, int penalty = 0,
                      int *up = NULL, int *down = NULL, int *left = NULL, int *right = NULL);
void GC_SetDataterms(GlobalWrapper *gw, int *data_positive, int *data_negative);
void GC_SetEdges(GlobalWrapper *gw, int *up = NULL, int *down = NULL, int *left = NULL, int *right = NULL);
void GC_SetPenalty(GlobalWrapper *gw, int p);
void GC_SetGraph(GlobalWrapper gw);
void GC_Optimize(GlobalWrapper gw, int *label);
void GC_End(GlobalWrapper *gw);

/* Additions over the reference API. */
void GC_SetOptions(GlobalWrapper gw, const gcfg_options *opt);
long long GC_GetFlow(GlobalWrapper gw);

#endif /* GRAPHCUT_H_ */
