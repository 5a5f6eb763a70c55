/* genomix_sizer.h — learned hash-table sizing for the MI355X graph build.
 *
 * None of these entry points has a counterpart in the reference: Genomix
 * never sizes its aggregation up front (its external sort spills and
 * VKmerList grows by 1.5x, VKmerList.java:135-139).  They serve callers
 * that want the device table sized before the insert: a small FP32 MLP,
 * trained on exact distinct-k-mer counts produced by the build itself,
 * reads a HyperLogLog sketch of the first 1% of the reads (plus k, read
 * length, coverage estimate, partitions) and predicts the distinct count
 * and the partition skew.  The graph output never depends on it: a too
 * small prediction falls back to grow-and-rebuild.
 *
 * All functions return GX_OK (0) or a negative GX_* code (genomix_gx.h).
 */
#ifndef GENOMIX_SIZER_H
#define GENOMIX_SIZER_H

#include <stddef.h>
#include <stdint.h>

#include "genomix_gx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gx_mlp gx_mlp;

/* Features of one read set (filled by the build at finish). */
#define GX_SIZER_FEATURES 72
/* Outputs: log(distinct / occurrences), log(max partition share * P). */
#define GX_SIZER_OUTPUTS 2
#define GX_SIZER_HLL_REGISTERS 4096

/* din -> tanh(hidden) -> tanh(hidden) -> dout, FP32, on one device;
 * parameters start at zero (set them with gx_mlp_set_params). */
int gx_mlp_create(int din, int hidden, int dout, int device, gx_mlp** out);
size_t gx_mlp_num_params(const gx_mlp* m);
/* Flat parameters, row-major [out][in]: W1, b1, W2, b2, W3, b3.  Setting
 * them also clears the Adam moments. */
int gx_mlp_set_params(gx_mlp* m, const float* params, size_t n);
int gx_mlp_get_params(gx_mlp* m, float* params, size_t n);
/* y[batch][dout] = f(x[batch][din]) (host arrays). */
int gx_mlp_forward(gx_mlp* m, const float* x, int batch, float* y);
/* Gradient of L = 1/(2B) sum (y - t)^2 with respect to the parameters
 * (grad may be NULL), and L. */
int gx_mlp_gradient(gx_mlp* m, const float* x, const float* t, int batch, float* grad, float* loss);
/* gx_mlp_gradient for callers that cannot allocate device memory (code
 * under HIP graph capture, tight training or serving loops): no counterpart
 * in the reference.  x [batch][din] and t [batch][dout] are DEVICE arrays;
 * grad (params floats) and loss (one float) are device outputs, either may
 * be NULL (the gradient then stays in the network's own buffer).  The
 * activations and deltas live in a fixed device array of the library (one
 * per GPU, 8.3 MiB, kept until the process ends), so batch must not exceed
 * 4096 samples (GX_EINVAL otherwise) and the call makes no device
 * allocation.  stream: a hipStream_t to enqueue on (the call then returns
 * without waiting), or NULL for the network's own stream, synchronized
 * before return.  Calls that overlap on one GPU share that array and race;
 * order them on one stream. */
int gx_mlp_gradient_noalloc(gx_mlp* m, const float* x, const float* t, int batch, float* grad, float* loss,
                            void* stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) on the batch; loss before the step. */
int gx_mlp_train_step(gx_mlp* m, const float* x, const float* t, int batch, float lr, float* loss);
const char* gx_mlp_last_error(const gx_mlp* m);
void gx_mlp_destroy(gx_mlp* m);

/* Attach a trained network to a build handle (NULL detaches).  With a
 * network attached and table_hint == 0, finish sizes the table to
 * ceil(1.5 * predicted distinct) slots; an overflow still grows the table
 * and rebuilds.  The network must live on the handle's device and outlive
 * its use. */
int gx_build_set_sizer(gx_build* b, gx_mlp* m);

/* After finish: the sizing features of the last build (n >= GX_SIZER_FEATURES),
 * the HLL registers of its sample (4096 bytes) and the network's prediction
 * [distinct k-mers, partition skew] (zeros when no network was attached);
 * each output may be NULL. */
int gx_build_features(gx_build* b, float* feat, int n, uint8_t* regs, double* prediction);

#ifdef __cplusplus
}
#endif

#endif /* GENOMIX_SIZER_H */
