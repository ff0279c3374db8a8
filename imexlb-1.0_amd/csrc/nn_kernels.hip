// nn_kernels.hip -- gfx950 kernels of the warm-start network: a small
// convolutional encoder-decoder trained and run on the GPU by these kernels
// (C ABI in include/imexlb_nn.h).  Tensors are FP32, NCHW, in device memory
// the caller owns; every launch goes to the given stream.
//
// conv3x3 (padding 1, stride 1): one block computes a 16x16 tile of pixels
// for 16 output channels; input channels stream through LDS 8 at a time
// (the tile with its one-pixel halo) together with the matching weights,
// stored [ci][tap][co] so one 128-bit LDS read feeds four FMAs.  The
// data gradient is the same kernel on the flipped, transposed weights; the
// weight gradient gives each thread one (co, ci) pair and its 9 taps over a
// tile, summed over the batch, then added to dW with atomics.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/imexlb_nn.h"

namespace imexlb_nn {

constexpr int TS = 16;   // pixel tile edge
constexpr int COB = 16;  // output channels per block
constexpr int CIB = 8;   // input channels per LDS chunk

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__global__ __launch_bounds__(256) void k_conv3(const float *__restrict__ in, const float *__restrict__ w,
                                               const float *__restrict__ bias, float *__restrict__ out, int N,
                                               int CI, int CO, int H, int W, int relu)
{
    __shared__ float tin[CIB][TS + 2][TS + 2];
    __shared__ float4 tw[CIB][9][COB / 4];
    const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
    const int x0 = blockIdx.x * TS, y0 = blockIdx.y * TS;
    const int ncob = (CO + COB - 1) / COB;
    const int n = blockIdx.z / ncob, co0 = (blockIdx.z % ncob) * COB;
    float acc[COB];
#pragma unroll
    for (int k = 0; k < COB; ++k) acc[k] = 0.0f;
    for (int ci0 = 0; ci0 < CI; ci0 += CIB) {
        for (int i = threadIdx.x; i < CIB * (TS + 2) * (TS + 2); i += 256) {
            const int c = i / ((TS + 2) * (TS + 2)), r = i % ((TS + 2) * (TS + 2));
            const int yy = y0 + r / (TS + 2) - 1, xx = x0 + r % (TS + 2) - 1;
            float v = 0.0f;
            if (ci0 + c < CI && yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = in[((int64_t(n) * CI + ci0 + c) * H + yy) * W + xx];
            tin[c][r / (TS + 2)][r % (TS + 2)] = v;
        }
        float *twf = &tw[0][0][0].x;
        for (int i = threadIdx.x; i < CIB * 9 * COB; i += 256) {
            const int c = i / (9 * COB), r = i % (9 * COB), k = r / COB, co = r % COB;
            float v = 0.0f;
            if (co0 + co < CO && ci0 + c < CI) v = w[(int64_t(co0 + co) * CI + ci0 + c) * 9 + k];
            twf[i] = v;
        }
        __syncthreads();
#pragma unroll 2
        for (int c = 0; c < CIB; ++c)
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const float v = tin[c][ty + k / 3][tx + k % 3];
#pragma unroll
                for (int q = 0; q < COB / 4; ++q) {
                    const float4 wv = tw[c][k][q];
                    acc[4 * q + 0] = fma_(wv.x, v, acc[4 * q + 0]);
                    acc[4 * q + 1] = fma_(wv.y, v, acc[4 * q + 1]);
                    acc[4 * q + 2] = fma_(wv.z, v, acc[4 * q + 2]);
                    acc[4 * q + 3] = fma_(wv.w, v, acc[4 * q + 3]);
                }
            }
        __syncthreads();
    }
    const int x = x0 + tx, y = y0 + ty;
    if (x >= W || y >= H) return;
#pragma unroll
    for (int co = 0; co < COB; ++co) {
        if (co0 + co >= CO) break;
        float v = acc[co] + (bias ? bias[co0 + co] : 0.0f);
        if (relu) v = v > 0.0f ? v : 0.0f;
        out[((int64_t(n) * CO + co0 + co) * H + y) * W + x] = v;
    }
}

// wt[ci][co][8-k] = w[co][ci][k]: the data gradient as a forward conv
__global__ void k_flip(const float *__restrict__ w, float *__restrict__ wt, int CI, int CO)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(CO) * CI * 9) return;
    const int k = int(i % 9), ci = int((i / 9) % CI), co = int(i / (9 * int64_t(CI)));
    wt[(int64_t(ci) * CO + co) * 9 + (8 - k)] = w[i];
}

constexpr int WCO = 16, WCI = 16;

__global__ __launch_bounds__(256) void k_conv3_dw(const float *__restrict__ in, const float *__restrict__ dout,
                                                  float *__restrict__ dw, int N, int CI, int CO, int H, int W)
{
    __shared__ float tin[WCI][TS + 2][TS + 2];
    __shared__ float tg[WCO][TS * TS];
    const int co = threadIdx.x / WCI, ci = threadIdx.x % WCI;
    const int ntx = (W + TS - 1) / TS;
    const int x0 = (blockIdx.x % ntx) * TS, y0 = (blockIdx.x / ntx) * TS;
    const int nci = (CI + WCI - 1) / WCI;
    const int co0 = (blockIdx.y / nci) * WCO, ci0 = (blockIdx.y % nci) * WCI;
    float acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = 0.0f;
    for (int n = 0; n < N; ++n) {
        for (int i = threadIdx.x; i < WCI * (TS + 2) * (TS + 2); i += 256) {
            const int c = i / ((TS + 2) * (TS + 2)), r = i % ((TS + 2) * (TS + 2));
            const int yy = y0 + r / (TS + 2) - 1, xx = x0 + r % (TS + 2) - 1;
            float v = 0.0f;
            if (ci0 + c < CI && yy >= 0 && yy < H && xx >= 0 && xx < W)
                v = in[((int64_t(n) * CI + ci0 + c) * H + yy) * W + xx];
            tin[c][r / (TS + 2)][r % (TS + 2)] = v;
        }
        for (int i = threadIdx.x; i < WCO * TS * TS; i += 256) {
            const int c = i / (TS * TS), p = i % (TS * TS);
            const int yy = y0 + p / TS, xx = x0 + p % TS;
            float v = 0.0f;
            if (co0 + c < CO && yy < H && xx < W) v = dout[((int64_t(n) * CO + co0 + c) * H + yy) * W + xx];
            tg[c][p] = v;
        }
        __syncthreads();
        for (int p = 0; p < TS * TS; ++p) {
            const float g = tg[co][p];
            const int py = p / TS, px = p % TS;
#pragma unroll
            for (int k = 0; k < 9; ++k) acc[k] = fma_(g, tin[ci][py + k / 3][px + k % 3], acc[k]);
        }
        __syncthreads();
    }
    if (co0 + co >= CO || ci0 + ci >= CI) return;
#pragma unroll
    for (int k = 0; k < 9; ++k) atomicAdd(&dw[(int64_t(co0 + co) * CI + ci0 + ci) * 9 + k], acc[k]);
}

// db[c] = sum of dout over batch and pixels (one block per channel)
__global__ __launch_bounds__(256) void k_bias_grad(const float *__restrict__ dout, float *__restrict__ db, int N,
                                                   int C, int64_t HW)
{
    __shared__ float red[256];
    const int c = blockIdx.x;
    float s = 0.0f;
    for (int n = 0; n < N; ++n)
        for (int64_t p = threadIdx.x; p < HW; p += 256) s += dout[(int64_t(n) * C + c) * HW + p];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = red[0];
}

__global__ void k_relu_bwd(const float *__restrict__ out, const float *__restrict__ dout, float *__restrict__ dz,
                           int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dz[i] = out[i] > 0.0f ? dout[i] : 0.0f;
}

// 2x2 average pool: in [N][C][2h][2w] -> out [N][C][h][w]
__global__ void k_pool2(const float *__restrict__ in, float *__restrict__ out, int64_t NC, int h, int w)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= NC * h * w) return;
    const int x = int(i % w), y = int((i / w) % h);
    const int64_t nc = i / (int64_t(h) * w);
    const float *p = in + (nc * 2 * h + 2 * y) * (2 * w) + 2 * x;
    out[i] = 0.25f * ((p[0] + p[1]) + (p[2 * w] + p[2 * w + 1]));
}

__global__ void k_pool2_bwd(const float *__restrict__ dout, float *__restrict__ din, int64_t NC, int h, int w)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // over din
    if (i >= NC * 4 * h * w) return;
    const int X = int(i % (2 * w)), Y = int((i / (2 * w)) % (2 * h));
    const int64_t nc = i / (int64_t(4) * h * w);
    din[i] = 0.25f * dout[(nc * h + Y / 2) * w + X / 2];
}

// nearest 2x upsample: in [N][C][h][w] -> channels [coff, coff+C) of out [N][Ctot][2h][2w]
__global__ void k_up2(const float *__restrict__ in, float *__restrict__ out, int N, int C, int h, int w, int Ctot,
                      int coff)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // over the C output channels
    const int64_t per = int64_t(C) * 4 * h * w;
    if (i >= N * per) return;
    const int n = int(i / per);
    const int64_t r = i % per;
    const int c = int(r / (int64_t(4) * h * w));
    const int X = int(r % (2 * w)), Y = int((r / (2 * w)) % (2 * h));
    out[((int64_t(n) * Ctot + coff + c) * 2 * h + Y) * 2 * w + X] = in[((int64_t(n) * C + c) * h + Y / 2) * w + X / 2];
}

__global__ void k_up2_bwd(const float *__restrict__ dout, float *__restrict__ din, int N, int C, int h, int w,
                          int Ctot, int coff)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;  // over din
    if (i >= int64_t(N) * C * h * w) return;
    const int x = int(i % w), y = int((i / w) % h);
    const int c = int((i / (int64_t(h) * w)) % C), n = int(i / (int64_t(C) * h * w));
    const float *p = dout + ((int64_t(n) * Ctot + coff + c) * 2 * h + 2 * y) * 2 * w + 2 * x;
    din[i] = (p[0] + p[1]) + (p[2 * w] + p[2 * w + 1]);
}

// dst[:, dcoff+c] (=|+=) src[:, scoff+c] for c < C
__global__ void k_copy_ch(const float *__restrict__ src, int sctot, int scoff, float *__restrict__ dst, int dctot,
                          int dcoff, int N, int C, int64_t HW, int accumulate)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int64_t(N) * C * HW) return;
    const int64_t p = i % HW;
    const int c = int((i / HW) % C), n = int(i / (HW * C));
    const float v = src[(int64_t(n) * sctot + scoff + c) * HW + p];
    float *d = dst + (int64_t(n) * dctot + dcoff + c) * HW + p;
    *d = accumulate ? *d + v : v;
}

// grad = 2 (pred - target) / n;  loss += sum (pred - target)^2 / n
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ pred, const float *__restrict__ target,
                                             float *__restrict__ grad, int64_t n, float *__restrict__ loss)
{
    __shared__ float red[256];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float e = 0.0f;
    if (i < n) {
        const float d = pred[i] - target[i];
        if (grad) grad[i] = 2.0f * d / float(n);
        e = d * d / float(n);
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (int(threadIdx.x) < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) atomicAdd(loss, red[0]);
}

__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (__builtin_sqrtf(vi / c2) + eps);
}

inline unsigned blocks(int64_t n) { return unsigned((n + 255) / 256); }

}  // namespace imexlb_nn

using namespace imexlb_nn;

#define NN_LAUNCH(...)                                                  \
    do {                                                                \
        hipLaunchKernelGGL(__VA_ARGS__);                                \
        return hipGetLastError() == hipSuccess ? 0 : -2;                \
    } while (0)

extern "C" {

int imexlb_nn_conv3x3(const float *in, const float *w, const float *bias, float *out, int N, int CI, int CO, int H,
                      int W, int relu, void *stream)
{
    if (!in || !w || !out || N <= 0 || CI <= 0 || CO <= 0 || H <= 0 || W <= 0) return -1;
    const dim3 grid(unsigned((W + TS - 1) / TS), unsigned((H + TS - 1) / TS), unsigned(N * ((CO + COB - 1) / COB)));
    NN_LAUNCH(k_conv3, grid, dim3(256), 0, (hipStream_t)stream, in, w, bias, out, N, CI, CO, H, W, relu);
}

int imexlb_nn_flip_weights(const float *w, float *wt, int CI, int CO, void *stream)
{
    if (!w || !wt || CI <= 0 || CO <= 0) return -1;
    NN_LAUNCH(k_flip, dim3(blocks(int64_t(CO) * CI * 9)), dim3(256), 0, (hipStream_t)stream, w, wt, CI, CO);
}

int imexlb_nn_conv3x3_dw(const float *in, const float *dout, float *dw, int N, int CI, int CO, int H, int W,
                         void *stream)
{
    if (!in || !dout || !dw || N <= 0 || CI <= 0 || CO <= 0 || H <= 0 || W <= 0) return -1;
    const dim3 grid(unsigned(((W + TS - 1) / TS) * ((H + TS - 1) / TS)),
                    unsigned(((CO + WCO - 1) / WCO) * ((CI + WCI - 1) / WCI)));
    NN_LAUNCH(k_conv3_dw, grid, dim3(256), 0, (hipStream_t)stream, in, dout, dw, N, CI, CO, H, W);
}

int imexlb_nn_bias_grad(const float *dout, float *db, int N, int C, int H, int W, void *stream)
{
    if (!dout || !db || N <= 0 || C <= 0 || H <= 0 || W <= 0) return -1;
    NN_LAUNCH(k_bias_grad, dim3(unsigned(C)), dim3(256), 0, (hipStream_t)stream, dout, db, N, C, int64_t(H) * W);
}

int imexlb_nn_relu_bwd(const float *out, const float *dout, float *dz, int64_t n, void *stream)
{
    if (!out || !dout || !dz || n <= 0) return -1;
    NN_LAUNCH(k_relu_bwd, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, out, dout, dz, n);
}

int imexlb_nn_pool2(const float *in, float *out, int N, int C, int h, int w, void *stream)
{
    if (!in || !out || N <= 0 || C <= 0 || h <= 0 || w <= 0) return -1;
    const int64_t NC = int64_t(N) * C;
    NN_LAUNCH(k_pool2, dim3(blocks(NC * h * w)), dim3(256), 0, (hipStream_t)stream, in, out, NC, h, w);
}

int imexlb_nn_pool2_bwd(const float *dout, float *din, int N, int C, int h, int w, void *stream)
{
    if (!dout || !din || N <= 0 || C <= 0 || h <= 0 || w <= 0) return -1;
    const int64_t NC = int64_t(N) * C;
    NN_LAUNCH(k_pool2_bwd, dim3(blocks(NC * 4 * h * w)), dim3(256), 0, (hipStream_t)stream, dout, din, NC, h, w);
}

int imexlb_nn_up2(const float *in, float *out, int N, int C, int h, int w, int Ctot, int coff, void *stream)
{
    if (!in || !out || N <= 0 || C <= 0 || h <= 0 || w <= 0 || coff < 0 || coff + C > Ctot) return -1;
    NN_LAUNCH(k_up2, dim3(blocks(int64_t(N) * C * 4 * h * w)), dim3(256), 0, (hipStream_t)stream, in, out, N, C, h,
              w, Ctot, coff);
}

int imexlb_nn_up2_bwd(const float *dout, float *din, int N, int C, int h, int w, int Ctot, int coff, void *stream)
{
    if (!dout || !din || N <= 0 || C <= 0 || h <= 0 || w <= 0 || coff < 0 || coff + C > Ctot) return -1;
    NN_LAUNCH(k_up2_bwd, dim3(blocks(int64_t(N) * C * h * w)), dim3(256), 0, (hipStream_t)stream, dout, din, N, C,
              h, w, Ctot, coff);
}

int imexlb_nn_copy_channels(const float *src, int sctot, int scoff, float *dst, int dctot, int dcoff, int N, int C,
                            int H, int W, int accumulate, void *stream)
{
    if (!src || !dst || N <= 0 || C <= 0 || H <= 0 || W <= 0 || scoff < 0 || dcoff < 0 || scoff + C > sctot ||
        dcoff + C > dctot)
        return -1;
    const int64_t HW = int64_t(H) * W;
    NN_LAUNCH(k_copy_ch, dim3(blocks(int64_t(N) * C * HW)), dim3(256), 0, (hipStream_t)stream, src, sctot, scoff,
              dst, dctot, dcoff, N, C, HW, accumulate);
}

int imexlb_nn_mse(const float *pred, const float *target, float *grad, int64_t n, float *loss, void *stream)
{
    if (!pred || !target || n <= 0) return -1;
    NN_LAUNCH(k_mse, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, pred, target, grad, n, loss);
}

int imexlb_nn_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps,
                   int step, void *stream)
{
    if (!w || !g || !m || !v || n <= 0 || step < 1) return -1;
    float c1 = 1.0f, c2 = 1.0f, p1 = 1.0f, p2 = 1.0f;
    for (int t = 0; t < step && t < 100000; ++t) {
        p1 *= b1;
        p2 *= b2;
    }
    c1 = 1.0f - p1;
    c2 = 1.0f - p2;
    NN_LAUNCH(k_adam, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

}  // extern "C"
