/* imexlb_nn.h -- kernels of the warm-start network (no counterpart in the
 * reference): a small convolutional encoder-decoder that predicts a
 * developed flow field from the body mask and Reynolds number, trained and
 * run on the GPU by these kernels.  Tensors are FP32, NCHW, in device memory
 * the caller owns; every call is asynchronous on `stream` (NULL: the default
 * stream).  Return 0, or -1 for a bad argument, -2 for a launch error. */
#ifndef IMEXLB_NN_H
#define IMEXLB_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* out[n,co] = act(bias[co] + sum_ci w[co,ci] (*) in[n,ci]), 3x3, padding 1;
 * w is [CO][CI][3][3]; bias may be NULL; relu: 1 applies max(0, .). */
int imexlb_nn_conv3x3(const float *in, const float *w, const float *bias, float *out, int N, int CI, int CO, int H,
                      int W, int relu, void *stream);
/* wt[ci][co][2-ky][2-kx] = w[co][ci][ky][kx]: with it, imexlb_nn_conv3x3 of
 * the output gradient gives the input gradient. */
int imexlb_nn_flip_weights(const float *w, float *wt, int CI, int CO, void *stream);
/* dw[co,ci,ky,kx] += sum over n, y, x of dout[n,co,y,x] in[n,ci,y+ky-1,x+kx-1] */
int imexlb_nn_conv3x3_dw(const float *in, const float *dout, float *dw, int N, int CI, int CO, int H, int W,
                         void *stream);
/* The same sum in an order that depends on the shapes only (results
 * repeat bit for bit): per-tile sums into work, then added to dw tile by
 * tile.  work holds imexlb_nn_conv3x3_dw_work(CI, CO, H, W) floats
 * (work_floats: its size; -1 if too small). */
int64_t imexlb_nn_conv3x3_dw_work(int CI, int CO, int H, int W);
int imexlb_nn_conv3x3_dw_ordered(const float *in, const float *dout, float *dw, float *work, int64_t work_floats,
                                 int N, int CI, int CO, int H, int W, void *stream);
/* imexlb_nn_conv3x3_dw_ordered without the workspace argument, for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * training loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  It allocates nothing: the per-tile sums go to a device array
 * of the library, 8,388,608 floats (32 MiB) on every GPU the library runs
 * on, held until the process ends.  Shapes whose
 * imexlb_nn_conv3x3_dw_work(CI, CO, H, W) exceeds that are rejected (-1;
 * imexlb_last_error says why); every layer of the 4-level 32..256-channel
 * network on 128 x 512 fits.  Results equal imexlb_nn_conv3x3_dw_ordered's
 * bit for bit.  Runs on the current device; calls that overlap on one GPU
 * (several streams or threads) share the array and race: serialise them. */
int imexlb_nn_conv3x3_dw_ordered_noalloc(const float *in, const float *dout, float *dw, int N, int CI, int CO,
                                         int H, int W, void *stream);
/* db[c] = sum of dout[:, c] */
int imexlb_nn_bias_grad(const float *dout, float *db, int N, int C, int H, int W, void *stream);
/* dz = dout where out > 0, else 0 (out: the ReLU output) */
int imexlb_nn_relu_bwd(const float *out, const float *dout, float *dz, int64_t n, void *stream);
/* 2x2 average pool, in [N][C][2h][2w] -> out [N][C][h][w], and its gradient */
int imexlb_nn_pool2(const float *in, float *out, int N, int C, int h, int w, void *stream);
int imexlb_nn_pool2_bwd(const float *dout, float *din, int N, int C, int h, int w, void *stream);
/* nearest 2x upsample of in [N][C][h][w] into channels [coff, coff+C) of
 * out [N][Ctot][2h][2w], and its gradient (din overwritten) */
int imexlb_nn_up2(const float *in, float *out, int N, int C, int h, int w, int Ctot, int coff, void *stream);
int imexlb_nn_up2_bwd(const float *dout, float *din, int N, int C, int h, int w, int Ctot, int coff, void *stream);
/* dst[:, dcoff+c] = src[:, scoff+c] (accumulate: +=), c < C; src has sctot
 * channels, dst dctot */
int imexlb_nn_copy_channels(const float *src, int sctot, int scoff, float *dst, int dctot, int dcoff, int N, int C,
                            int H, int W, int accumulate, void *stream);
/* mean squared error of n values: grad = 2 (pred - target) / n (grad may be
 * NULL); *loss += the mean (loss may be NULL) */
int imexlb_nn_mse(const float *pred, const float *target, float *grad, int64_t n, float *loss, void *stream);
/* Adam step `step` (1-based) on n weights */
int imexlb_nn_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                   int step, void *stream);

/* BF16 inference.  Activations are NHWC bf16 (uint16_t bit patterns),
 * weights packed by imexlb_nn_pack_weights_bf16; products accumulate in
 * FP32 on the matrix cores; values are rounded to bf16 to nearest even.
 * Buffers given as uint16_t pointers must be 16-byte aligned. */
/* x [N][C][H][W] FP32 -> y [N][H][W][CP] bf16 (channels C..CP-1 zero) */
int imexlb_nn_pack_input_bf16(const float *x, uint16_t *y, int N, int C, int CP, int H, int W, void *stream);
/* w [CO][CI][3][3] FP32 -> wp [CIP/8][9][COP][8] bf16, zero-padded;
 * CIP a multiple of 8, COP >= CO */
int imexlb_nn_pack_weights_bf16(const float *w, uint16_t *wp, int CI, int CO, int CIP, int COP, void *stream);
/* 3x3 convolution, padding 1, of in [N][H][W][CIP] (CIP a multiple of 16)
 * with packed wp (COP a multiple of 32): output channel co < CO of pixel
 * (n, y, x) = act(bias[co] + sum), written as bf16 to
 * out[((n H + y) W + x) ostride + ooff + co] (ooff + CO <= ostride), or,
 * with f32_nchw = 1, as FP32 to out[((n CO + co) H + y) W + x].  bias may
 * be NULL; relu: 1 applies max(0, .); N <= 65535. */
int imexlb_nn_conv3x3_bf16(const uint16_t *in, const uint16_t *wp, const float *bias, void *out, int N, int CIP,
                           int CO, int COP, int H, int W, int ostride, int ooff, int relu, int f32_nchw, void *stream);
/* 2x2 average pool of channels [coff, coff+C) of in [N][2h][2w][cstride]
 * into out [N][h][w][C] (C, cstride, coff multiples of 8) */
int imexlb_nn_pool2_bf16(const uint16_t *in, uint16_t *out, int N, int C, int h, int w, int cstride, int coff,
                         void *stream);
/* nearest 2x upsample of in [N][h][w][C] into channels [coff, coff+C) of
 * out [N][2h][2w][Ctot] (C, Ctot, coff multiples of 8) */
int imexlb_nn_up2_bf16(const uint16_t *in, uint16_t *out, int N, int C, int h, int w, int Ctot, int coff,
                       void *stream);

#ifdef __cplusplus
}
#endif
#endif /* IMEXLB_NN_H */
