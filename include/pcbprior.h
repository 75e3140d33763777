/* pcbprior.h - C ABI of the learned corridor / congestion prior: a small conv3x3 CNN trained and run on the MI355X
 * with BF16 matrix-core (MFMA) kernels that accumulate in FP32.
 *
 * It replaces the ONNX Runtime call of pcbroute's learned prior (src/mlinfer.cpp:834-856: x (n,12,h,w) ->
 * Session::Run -> (n,4,h,w) probabilities), with a network of its own: `depth` conv3x3 layers of `width` channels
 * (ReLU between them), in_ch input channels per pixel, out_ch corridor logits per pixel (one per router layer).
 * Pixels are the router grid downsampled 4x. Tensors cross the boundary as host float arrays, pixel-major (NHWC):
 * x[((i * h + y) * w + x) * in_ch + c]. Spatial sizes: w a multiple of 64 (pad the crop).
 *
 * The prior enters a search only through pcbgpu_query's prior map (an integer cost bias per prior value, clamped to
 * [0, 8], and a corridor threshold), and the plain search stays the fallback (src/router.cpp:1804).
 */
#ifndef PCBPRIOR_H
#define PCBPRIOR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCBPRIOR_MAX_CH 32  /* channels a layer carries (inputs and outputs are padded to it) */

typedef struct pcbprior pcbprior;

/* A network of `depth` conv3x3 layers (in_ch -> width -> ... -> out_ch), weights drawn from `seed`. NULL on error. */
pcbprior* pcbprior_create(int in_ch, int width, int depth, int out_ch, uint64_t seed);
void pcbprior_destroy(pcbprior* p);

/* Parameter count (FP32 master weights and biases, layer by layer: w[co][ky][kx][ci] then b[co]). */
size_t pcbprior_num_params(const pcbprior* p);
int pcbprior_get_params(const pcbprior* p, float* out);
int pcbprior_set_params(pcbprior* p, const float* in);

/* Forward pass: per-pixel corridor probabilities (sigmoid of the logits), out (n, h, w, out_ch). */
int pcbprior_predict(pcbprior* p, const float* x, int n, int h, int w, float* prob);
/* The same over a batch of images of different sizes padded to h x w: image k's valid extent is hw[2k] rows by
 * hw[2k + 1] columns (a multiple of 64), its inputs zero outside it. Every output inside an image's extent equals
 * pcbprior_predict on that image alone at its extent; outputs outside are sigmoid(0). hw NULL: all valid. */
int pcbprior_predict_valid(pcbprior* p, const float* x, int n, int h, int w, const int32_t* hw, float* prob);
/* Logits instead of probabilities. */
int pcbprior_logits(pcbprior* p, const float* x, int n, int h, int w, float* logits);

/* Loss and gradient: mean binary cross-entropy of the logits against target (n, h, w, out_ch) in [0, 1] over the
 * pixels where mask (n, h, w) is nonzero; grads gets d loss / d params in pcbprior_get_params's order. */
int pcbprior_loss_grad(pcbprior* p, const float* x, const float* target, const float* mask, int n, int h, int w,
                       float* loss, float* grads);

/* One Adam step on a batch (the loss is the one above). */
int pcbprior_train_step(pcbprior* p, const float* x, const float* target, const float* mask, int n, int h, int w,
                        float lr, float* loss);

/* ---- a training set resident in HBM (no counterpart in the reference): the samples go over once, each step
 * gathers its batch on the device, so training copies nothing per step. x, target, mask as above with n samples of
 * h x w pixels. */
int pcbprior_dataset(pcbprior* p, const float* x, const float* target, const float* mask, int n, int h, int w);
/* one Adam step on samples idx[0..batch) of the resident set; loss may be NULL */
int pcbprior_train_resident(pcbprior* p, const int32_t* idx, int batch, float lr, float* loss);
/* the forward pass over samples idx[0..batch); logits (batch, h, w, out_ch) may be NULL */
int pcbprior_forward_resident(pcbprior* p, const int32_t* idx, int batch, float* logits);

/* Kernel accounting: while enabled, every conv3x3 launch (forward and input gradient) and weight-gradient launch
 * is bracketed by HIP events on the network's stream; FLOPs count real channels (2 x pixels x cout x 9 x cin). */
typedef struct pcbprior_stats {
    uint64_t conv_launches, wgrad_launches;
    double conv_ms, wgrad_ms;
    double conv_flop, wgrad_flop;
} pcbprior_stats;
void pcbprior_stats_enable(pcbprior* p, int on);
void pcbprior_stats_read(pcbprior* p, pcbprior_stats* s, int reset);

/* One 16x16x32 BF16 matrix-core product (the shape the conv kernels use), for the lane-layout test: a (16x32),
 * b (32x16) row-major floats rounded to BF16, d (16x16) = a @ b in FP32. */
int pcbprior_mfma_check(const float* a, const float* b, float* d);

const char* pcbprior_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
