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
//
// Inference in BF16: activations NHWC bf16, weights packed
// [CI/8][tap][CO][8 ci] bf16, products on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation), bias, ReLU and the
// rounding to bf16 (round to nearest even) fused into the store.  One block
// computes 8x16 pixels for 32 output channels, one 32-pixel row pair per
// wave; the input tile with its halo and the block's weights go through LDS
// 16 channels at a time (15 KB), stored in 16-byte groups of 8 channels so
// that every lane's operand is one 128-bit LDS read and the 32 lanes of a
// half-wave read consecutive groups.  The small blocks keep up to 8 of them
// on a CU, which hides the staging better than larger blocks with more
// products per chunk (measured: 32- or 64-channel chunks and 64-channel
// blocks run 15-20 % slower).  Skip connections cost no copy: an
// encoder convolution writes into its channel slice of the decoder's
// concatenated input (output channel stride and offset).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/imexlb_nn.h"
#include "context.hpp"

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

// partial == nullptr: add to dw with atomics; otherwise write the block's
// sums to partial[blockIdx.x][co][ci][9] for k_dw_reduce (fixed order).
__global__ __launch_bounds__(256) void k_conv3_dw(const float *__restrict__ in, const float *__restrict__ dout,
                                                  float *__restrict__ dw, int N, int CI, int CO, int H, int W,
                                                  float *__restrict__ partial)
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
    const int64_t o = (int64_t(co0 + co) * CI + ci0 + ci) * 9;
    if (partial) {
        float *pb = partial + int64_t(blockIdx.x) * CO * CI * 9 + o;
#pragma unroll
        for (int k = 0; k < 9; ++k) pb[k] = acc[k];
        return;
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) atomicAdd(&dw[o + k], acc[k]);
}

// dw[i] += sum over tiles b = 0..nb-1 of partial[b][i], in that order
__global__ void k_dw_reduce(const float *__restrict__ partial, float *__restrict__ dw, int64_t nb, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int64_t b = 0; b < nb; ++b) s += partial[b * n + i];
    dw[i] += s;
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

// ------------------------------------------------------------ BF16 inference
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> bf16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t to_bf16(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return uint16_t((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint32_t h) { return __uint_as_float(h << 16); }

constexpr int PR = 8, PC = 16;  // pixel tile of the BF16 convolution: 8 rows x 16 columns

// KC input channels per LDS chunk (16 or 32), BN output channels per block
// (32 or 64).  in: [N][H][W][CIP] bf16 as 16-byte groups; wp: [CIP/8][9][COP]
// groups of 8 input channels.  Output channel co < CO of pixel (n, y, x):
// bf16 at out[((n H + y) W + x) ostride + ooff + co], or with f32_nchw FP32
// at out[((n CO + co) H + y) W + x].
template <int KC, int BN>
__global__ __launch_bounds__(256) void k_conv3_bf16(const uint4 *__restrict__ in, const uint4 *__restrict__ wp,
                                                    const float *__restrict__ bias, void *__restrict__ out, int H,
                                                    int W, int CIP, int CO, int COP, int ostride, int ooff, int relu,
                                                    int f32_nchw)
{
    constexpr int G = KC / 8;  // channel groups per chunk
    constexpr int NT = BN / 32;
    __shared__ uint4 s_in[G][PR + 2][PC + 2];
    __shared__ uint4 s_w[9][G][BN];
    const int lane = int(threadIdx.x) % 64, wave = int(threadIdx.x) / 64;
    const int ntx = (W + PC - 1) / PC;
    const int x0 = int(blockIdx.x) % ntx * PC, y0 = int(blockIdx.x) / ntx * PR;
    const int co0 = int(blockIdx.y) * BN, n = int(blockIdx.z);
    const int cg = CIP / 8;
    const int m = lane % 32, hk = lane / 32;        // operand row (pixel / channel) and K half
    const int pr = wave * 2 + m / 16, pc = m % 16;  // this lane's pixel in the tile
    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[t][j] = 0.0f;
    const int64_t img = int64_t(n) * H;
    for (int c0 = 0; c0 < CIP; c0 += KC) {
        for (int e = int(threadIdx.x); e < (PR + 2) * (PC + 2) * G; e += 256) {
            const int g = e % G, p = e / G;
            const int r = p / (PC + 2), c = p % (PC + 2);
            const int yy = y0 + r - 1, xx = x0 + c - 1;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) v = in[((img + yy) * W + xx) * cg + c0 / 8 + g];
            s_in[g][r][c] = v;
        }
        for (int e = int(threadIdx.x); e < 9 * G * BN; e += 256) {
            const int co = e % BN, t = (e / BN) % 9, g = e / (9 * BN);
            s_w[t][g][co] = wp[(int64_t(c0 / 8 + g) * 9 + t) * COP + co0 + co];
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int s = 0; s < KC / 16; ++s) {
                const bf16x8 a = __builtin_bit_cast(bf16x8, s_in[2 * s + hk][pr + t / 3][pc + t % 3]);
#pragma unroll
                for (int q = 0; q < NT; ++q) {
                    const bf16x8 b = __builtin_bit_cast(bf16x8, s_w[t][2 * s + hk][q * 32 + m]);
                    acc[q] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[q], 0, 0, 0);
                }
            }
        __syncthreads();
    }
    // lane holds D[pixel 8 (j/4) + 4 hk + j%4][channel m] of each 32x32 tile
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const int co = co0 + q * 32 + m;
        if (co >= CO) continue;
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int px = 8 * (j / 4) + 4 * hk + j % 4;
            const int y = y0 + wave * 2 + px / 16, x = x0 + px % 16;
            if (y >= H || x >= W) continue;
            float v = acc[q][j] + bv;
            if (relu) v = v > 0.0f ? v : 0.0f;
            if (f32_nchw)
                static_cast<float *>(out)[((int64_t(n) * CO + co) * H + y) * W + x] = v;
            else
                static_cast<uint16_t *>(out)[((img + y) * W + x) * ostride + ooff + co] = to_bf16(v);
        }
    }
}

// x [N][C][H][W] FP32 -> y [N][H][W][CP] bf16, channels C..CP-1 zero
__global__ void k_pack_in(const float *__restrict__ x, uint16_t *__restrict__ y, int C, int CP, int H, int W,
                          int64_t total)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int c = int(i % CP);
    const int64_t p = i / CP;
    const int64_t hw = int64_t(H) * W;
    const int64_t n = p / hw, r = p % hw;
    y[i] = c < C ? to_bf16(x[(n * C + c) * hw + r]) : uint16_t(0);
}

// w [CO][CI][3][3] FP32 -> wp [CIP/8][9][COP][8] bf16, zero outside CO x CI
__global__ void k_pack_w(const float *__restrict__ w, uint16_t *__restrict__ wp, int CI, int CO, int COP,
                         int64_t total)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int j = int(i % 8), co = int((i / 8) % COP), t = int((i / (8 * int64_t(COP))) % 9);
    const int ci = int(i / (72 * int64_t(COP))) * 8 + j;
    wp[i] = (co < CO && ci < CI) ? to_bf16(w[(int64_t(co) * CI + ci) * 9 + t]) : uint16_t(0);
}

// 2x2 average pool of channels [coff, coff + 8 C8) of in [N][2h][2w][8 cs8]
// into out [N][h][w][8 C8]; one thread per 8 channels
__global__ void k_pool2_bf16(const uint4 *__restrict__ in, uint4 *__restrict__ out, int h, int w, int cs8, int coff8,
                             int C8, int64_t total)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int g = int(i % C8);
    const int64_t p = i / C8;
    const int x = int(p % w), y = int((p / w) % h);
    const int64_t n = p / (int64_t(w) * h);
    const int64_t b = ((n * 2 * h + 2 * y) * 2 * w + 2 * x) * cs8 + coff8 + g;
    const uint4 v[4] = {in[b], in[b + cs8], in[b + 2 * int64_t(w) * cs8], in[b + (2 * int64_t(w) + 1) * cs8]};
    uint32_t o[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint32_t r = 0;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            float s[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t word = k == 0 ? v[q].x : k == 1 ? v[q].y : k == 2 ? v[q].z : v[q].w;
                s[q] = from_bf16(half ? word >> 16 : word & 0xffffu);
            }
            r |= uint32_t(to_bf16(0.25f * ((s[0] + s[1]) + (s[2] + s[3])))) << (16 * half);
        }
        o[k] = r;
    }
    out[i] = make_uint4(o[0], o[1], o[2], o[3]);
}

// nearest 2x upsample of in [N][h][w][8 C8] into channels [coff, coff + 8 C8)
// of out [N][2h][2w][8 ct8]
__global__ void k_up2_bf16(const uint4 *__restrict__ in, uint4 *__restrict__ out, int h, int w, int ct8, int coff8,
                           int C8, int64_t total)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int g = int(i % C8);
    const int64_t p = i / C8;
    const int X = int(p % (2 * w)), Y = int((p / (2 * w)) % (2 * h));
    const int64_t n = p / (int64_t(4) * w * h);
    out[p * ct8 + coff8 + g] = in[((n * h + Y / 2) * w + X / 2) * C8 + g];
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
    NN_LAUNCH(k_conv3_dw, grid, dim3(256), 0, (hipStream_t)stream, in, dout, dw, N, CI, CO, H, W, nullptr);
}

int64_t imexlb_nn_conv3x3_dw_work(int CI, int CO, int H, int W)
{
    if (CI <= 0 || CO <= 0 || H <= 0 || W <= 0) return -1;
    return int64_t((W + TS - 1) / TS) * ((H + TS - 1) / TS) * CO * CI * 9;
}

int imexlb_nn_conv3x3_dw_ordered(const float *in, const float *dout, float *dw, float *work, int64_t work_floats,
                                 int N, int CI, int CO, int H, int W, void *stream)
{
    if (!in || !dout || !dw || !work || N <= 0 || CI <= 0 || CO <= 0 || H <= 0 || W <= 0) return -1;
    const int64_t tiles = int64_t((W + TS - 1) / TS) * ((H + TS - 1) / TS);
    const int64_t n = int64_t(CO) * CI * 9;
    if (work_floats < tiles * n) return -1;
    const dim3 grid(unsigned(tiles), unsigned(((CO + WCO - 1) / WCO) * ((CI + WCI - 1) / WCI)));
    hipLaunchKernelGGL(k_conv3_dw, grid, dim3(256), 0, (hipStream_t)stream, in, dout, dw, N, CI, CO, H, W, work);
    if (hipGetLastError() != hipSuccess) return -2;
    NN_LAUNCH(k_dw_reduce, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, work, dw, tiles, n);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The ordered weight gradient without a caller workspace: the per-tile sums
// go to `scratch`, a device array of the library that every GPU it runs on
// holds until the process ends.  Host code reaches it only through the
// address hipGetSymbolAddress gives on each device (the array's name is a
// host-side stand-in).  Nothing in it needs to start at zero: k_conv3_dw
// writes every partial that k_dw_reduce then reads.  Calls that overlap on
// one GPU share the array and race.
constexpr int64_t kScratchFloats = int64_t(8) << 20;  // 8 Mi floats, 32 MiB
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
float *scratch_dev[kMaxDevices] = {};

// this device's copy of scratch (the caller's current device), looked up on
// the first call on the device; nullptr with the error text set otherwise
float *scratch_here()
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        imexlb::fail(IMEXLB_E_HIP, std::string("no usable current device: ") + hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            imexlb::fail(IMEXLB_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        scratch_dev[dev] = static_cast<float *>(p);
    }
    return scratch_dev[dev];
}

}  // namespace

extern "C" {

int imexlb_nn_conv3x3_dw_ordered_noalloc(const float *in, const float *dout, float *dw, int N, int CI, int CO,
                                         int H, int W, void *stream)
{
    if (!in || !dout || !dw || N <= 0 || CI <= 0 || CO <= 0 || H <= 0 || W <= 0) return -1;
    const int64_t tiles = int64_t((W + TS - 1) / TS) * ((H + TS - 1) / TS);
    const int64_t n = int64_t(CO) * CI * 9;
    if (tiles * n > kScratchFloats) {
        imexlb::fail(IMEXLB_E_ARG, "weight gradient needs " + std::to_string(tiles * n) +
                                       " floats of partial sums; the library's scratch holds " +
                                       std::to_string(kScratchFloats));
        return -1;
    }
    float *work = scratch_here();
    if (!work) return -2;
    const dim3 grid(unsigned(tiles), unsigned(((CO + WCO - 1) / WCO) * ((CI + WCI - 1) / WCI)));
    hipLaunchKernelGGL(k_conv3_dw, grid, dim3(256), 0, (hipStream_t)stream, in, dout, dw, N, CI, CO, H, W, work);
    if (hipGetLastError() != hipSuccess) return -2;
    NN_LAUNCH(k_dw_reduce, dim3(blocks(n)), dim3(256), 0, (hipStream_t)stream, work, dw, tiles, n);
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

int imexlb_nn_pack_input_bf16(const float *x, uint16_t *y, int N, int C, int CP, int H, int W, void *stream)
{
    if (!x || !y || N <= 0 || C <= 0 || CP < C || H <= 0 || W <= 0) return -1;
    const int64_t total = int64_t(N) * H * W * CP;
    NN_LAUNCH(k_pack_in, dim3(blocks(total)), dim3(256), 0, (hipStream_t)stream, x, y, C, CP, H, W, total);
}

int imexlb_nn_pack_weights_bf16(const float *w, uint16_t *wp, int CI, int CO, int CIP, int COP, void *stream)
{
    if (!w || !wp || CI <= 0 || CO <= 0 || CIP < CI || CIP % 8 || COP < CO) return -1;
    const int64_t total = int64_t(CIP) * 9 * COP;
    NN_LAUNCH(k_pack_w, dim3(blocks(total)), dim3(256), 0, (hipStream_t)stream, w, wp, CI, CO, COP, total);
}

int imexlb_nn_conv3x3_bf16(const uint16_t *in, const uint16_t *wp, const float *bias, void *out, int N, int CIP,
                           int CO, int COP, int H, int W, int ostride, int ooff, int relu, int f32_nchw, void *stream)
{
    if (!in || !wp || !out || N <= 0 || N > 65535 || H <= 0 || W <= 0 || CIP <= 0 || CIP % 16 || CO <= 0 ||
        COP < CO || COP % 32)
        return -1;
    if (!f32_nchw && (ooff < 0 || ooff + CO > ostride)) return -1;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(wp)) & 15) return -1;
    const int64_t tiles = int64_t((W + PC - 1) / PC) * ((H + PR - 1) / PR);
    if (tiles > 0x7fffffff) return -1;
    const dim3 grid(unsigned(tiles), unsigned(COP / 32), unsigned(N));
    hipLaunchKernelGGL((k_conv3_bf16<16, 32>), grid, dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const uint4 *>(in), reinterpret_cast<const uint4 *>(wp), bias, out, H, W, CIP,
                       CO, COP, ostride, ooff, relu, f32_nchw);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int imexlb_nn_pool2_bf16(const uint16_t *in, uint16_t *out, int N, int C, int h, int w, int cstride, int coff,
                         void *stream)
{
    if (!in || !out || N <= 0 || C <= 0 || h <= 0 || w <= 0 || C % 8 || cstride % 8 || coff % 8 || coff < 0 ||
        coff + C > cstride)
        return -1;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) return -1;
    const int64_t total = int64_t(N) * h * w * (C / 8);
    NN_LAUNCH(k_pool2_bf16, dim3(blocks(total)), dim3(256), 0, (hipStream_t)stream,
              reinterpret_cast<const uint4 *>(in), reinterpret_cast<uint4 *>(out), h, w, cstride / 8, coff / 8, C / 8,
              total);
}

int imexlb_nn_up2_bf16(const uint16_t *in, uint16_t *out, int N, int C, int h, int w, int Ctot, int coff,
                       void *stream)
{
    if (!in || !out || N <= 0 || C <= 0 || h <= 0 || w <= 0 || C % 8 || Ctot % 8 || coff % 8 || coff < 0 ||
        coff + C > Ctot)
        return -1;
    if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) return -1;
    const int64_t total = int64_t(N) * 4 * h * w * (C / 8);
    NN_LAUNCH(k_up2_bf16, dim3(blocks(total)), dim3(256), 0, (hipStream_t)stream,
              reinterpret_cast<const uint4 *>(in), reinterpret_cast<uint4 *>(out), h, w, Ctot / 8, coff / 8, C / 8,
              total);
}

}  // extern "C"
