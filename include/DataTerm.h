/*
 * DataTerm.h -- drop-in C++ host API of the reference colour data term
 * (DataTerm.h:7-38, DataTerm.cpp:12-165), backed by libgcfg.so.
 *
 * Same namespace, function names and single-instance semantics. Two colour
 * models (setModel):
 *  - GCFG_DT_MODEL_KDE (default, the reference's model): the Gaussian kernel
 *    density of DataTermKernels.cu:19-150: selectOnMap(x, y, sigma) adds the
 *    colours of the 5x5 square at (x..x+4, y..y+4) to the density,
 *    setSigmaGaussian sets its width, and pos = (int)(300000 p),
 *    neg = (int)(300000 (1 - p));
 *  - GCFG_DT_MODEL_HISTOGRAM (opt-in): 16 bins per channel and class, +1
 *    smoothing, pos/neg = llround(64 * -ln p), contrast n-links with lambda 50;
 *    selectOnMap marks the same square as foreground seeds, selectRect adds a
 *    rectangle of either class.
 *  - getDataTermsMixture() computes the t-links (and contrast n-links);
 *  - getPos()/getNeg() return device int* t-links for GC_SetDataterms,
 *    getUp/Down/Left/Right device n-links for GC_SetEdges;
 *  - selectOnMap marks the square in the selection map, and
 *    setSelection/drawSelection keep the reference's display behaviour
 *    (selected pixels painted (0, 0, 255) in BGR order, DataTermKernels.cu:96-113).
 * Errors print a message and abort, as the reference's CUDA_SAFE_CALL does.
 */
#ifndef DATATERM_H
#define DATATERM_H

#include "gcfg.h"

namespace DataTerm {
int *getPos();
int *getNeg();
void init_data_term(int w, int h);
void free_data_term();
void setImage(unsigned char *h_img);
void setSelection(int *h_select);
void reset();
void selectOnMap(int x, int y, int sigma);
void drawSelection(unsigned char *inp, unsigned char *out);
void getDataTermsMixture();
void setSigmaGaussian(int sigma);

/* Additions over the reference API. */
int *getUp();
int *getDown();
int *getLeft();
int *getRight();
void selectRect(int x0, int y0, int w, int h, int cls); /* cls: 1 foreground, 2 background */
void setLambda(int lambda);
void setModel(int model); /* GCFG_DT_MODEL_KDE (default) or GCFG_DT_MODEL_HISTOGRAM */
}  // namespace DataTerm

#endif
