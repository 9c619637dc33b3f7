/*
 * DataTerm.h -- drop-in C++ host API of the reference colour data term
 * (DataTerm.h:7-38, DataTerm.cpp:12-165), backed by libgcfg.so.
 *
 * Same namespace, function names and single-instance semantics. The colour
 * model behind it is the histogram model of gcfg.h (16 bins per channel and
 * class, +1 smoothing, pos/neg = llround(64 * -ln p), contrast n-links with
 * lambda 50) instead of the reference's 256^3 Gaussian kernel density:
 *  - selectOnMap(x, y, sigma) marks the 5x5 square at (x..x+4, y..y+4) as
 *    foreground seeds (DataTerm.cpp:105-127 selects the same square);
 *    selectRect adds a rectangle of either class;
 *  - getDataTermsMixture() computes histograms, t-links and n-links;
 *  - getPos()/getNeg() return device int* t-links for GC_SetDataterms,
 *    getUp/Down/Left/Right device n-links for GC_SetEdges;
 *  - setSigmaGaussian is accepted and ignored (no kernel density);
 *  - setSelection/drawSelection keep the reference's display behaviour
 *    (selected pixels painted (0, 0, 255) in BGR order, DataTermKernels.cu:96-113).
 * Errors print a message and abort, as the reference's CUDA_SAFE_CALL does.
 */
#ifndef DATATERM_H
#define DATATERM_H

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
}  // namespace DataTerm

#endif
