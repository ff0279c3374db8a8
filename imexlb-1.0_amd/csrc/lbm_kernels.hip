// lbm_kernels.hip -- gfx950 kernels of the lattice Boltzmann time step.
//
// Device layout: populations live on a padded grid with ONE ghost layer,
// padded cell c = o0 + x + Px*y + PxPy*z (x in [-1, nx+1] ...), stored in
// tiles of IMEXLB_TILE consecutive cells with the Q directions of a tile one
// after another: slot(q, c) = ((c/T)*Q + q)*T + c%T (T = 0 selects plain
// planes, slot = q*plane + c).
// Interior fluid nodes hold post-collision values; interior solid nodes
// hold 0 (the reference never writes them, lbm.f90:57); the ghost layer and
// the body cells the reference rewrites (its b_user list) take their values
// from a resolved table that reproduces PassF plus the five copy phases of
// BoundaryCondition (lbm.f90:96-155) slot for slot -- either as its own fill
// kernel (k_fill) or written by the fused step that produces the buffer
// (K_PUSH, the table grouped by source node).
//
// One thread per padded cell (threads on ghost cells leave), x fastest, so a
// wave covers one tile.  In FP64 form every kernel keeps the reference's
// operation order (lbm.f90:56-74, 163-199) and the library is compiled
// without floating-point contraction, so results are the reference's
// IEEE-754 results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "kernels.hpp"

// Build-time tuning of the step kernel (defaults measured best on MI355X):
// threads per block; non-temporal (streaming) stores for the post-collision
// output -- with the tiled layout they save 7-14 %, with planes they cost
// 17 %; non-temporal loads (cost 8 %).
#ifndef IMEXLB_STEP_BLOCK
#define IMEXLB_STEP_BLOCK 256
#endif
#ifndef IMEXLB_STEP_NT
#define IMEXLB_STEP_NT 1
#endif
#ifndef IMEXLB_STEP_NTLOAD
#define IMEXLB_STEP_NTLOAD 0
#endif
// minimum waves per SIMD the step kernel is compiled for (0: compiler's
// choice); it caps the registers per thread at 512/waves
#ifndef IMEXLB_STEP_WAVES
#define IMEXLB_STEP_WAVES 0
#endif
// cells per thread of the fused collide step (1, 2 or 4)
#ifndef IMEXLB_STEP_CPT
#define IMEXLB_STEP_CPT 1
#endif

namespace imexlb {

// ---------------------------------------------------------------------------
// lattice tables (reference numbering, initialization.f90:112-123 and
// initialization_3d.f90:108-122)
template <int D>
struct Lat {
    static constexpr int Q = D == 2 ? 9 : 27;
    __host__ __device__ static constexpr int ex(int q) { return (q % 9) / 3 - 1; }
    __host__ __device__ static constexpr int ey(int q) { return q % 3 - 1; }
    __host__ __device__ static constexpr int ez(int q) { return D == 3 ? q / 9 - 1 : 0; }
    __host__ __device__ static constexpr int norm1(int q)
    {
        return (ex(q) < 0 ? -ex(q) : ex(q)) + (ey(q) < 0 ? -ey(q) : ey(q)) + (ez(q) < 0 ? -ez(q) : ez(q));
    }
    // 2D: 4*(1-0.75ex^2)(1-0.75ey^2) is exact (4, 1, 0.25), then /9.
    // 3D: 8/4^n/27.  Both divisions correctly rounded, as in the reference.
    __host__ __device__ static constexpr double w(int q)
    {
        return D == 2 ? (norm1(q) == 0 ? 4.0 / 9.0 : norm1(q) == 1 ? 1.0 / 9.0 : 0.25 / 9.0)
                      : (norm1(q) == 0 ? 8.0 / 27.0
                                       : norm1(q) == 1 ? 2.0 / 27.0 : norm1(q) == 2 ? 0.5 / 27.0 : 0.125 / 27.0);
    }
};

// ---------------------------------------------------------------------------
// storage forms of a population value
struct bf16_t {
    uint16_t bits;
};
// BF16 storage with stochastic rounding (IMEXLB_BF16SR_F32)
struct bf16sr_t {
    uint16_t bits;
};

// st(v, word): the stored form of v; word (a uniformly distributed 32-bit
// value per slot and step, sr_word below) only feeds the stochastic rounding
template <class S>
struct Store;
template <>
struct Store<double> {
    __device__ static double ld(double v) { return v; }
    template <class C>
    __device__ static double st(C v, uint32_t = 0u) { return double(v); }
};
template <>
struct Store<float> {
    __device__ static float ld(float v) { return v; }
    template <class C>
    __device__ static float st(C v, uint32_t = 0u) { return float(v); }
};
template <>
struct Store<_Float16> {
    __device__ static float ld(_Float16 v) { return float(v); }
    template <class C>
    __device__ static _Float16 st(C v, uint32_t = 0u) { return _Float16(float(v)); }  // round to nearest even
};
// two BF16 values in one register, the operand of the conversion builtins
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
template <>
struct Store<bf16_t> {
    __device__ static float ld(bf16_t v) { return __uint_as_float(uint32_t(v.bits) << 16); }
    // round to nearest even in one instruction (gfx950's v_cvt_pk_bf16_f32;
    // a NaN stays a quiet NaN)
    template <class C>
    __device__ static bf16_t st(C v, uint32_t = 0u)
    {
        return bf16_t{__builtin_bit_cast(uint16_t, static_cast<__bf16>(float(v)))};
    }
    // The same rounding through the stochastic-rounding instruction, whose
    // word is then the round-to-nearest-even increment 0x7fff + lsb in its
    // high half: three instructions, equal to st() for every input but NaN
    // payloads.  The D3Q27 step with LDS copies is scheduled 15 % slower
    // around the one-instruction form.
    template <class C>
    __device__ static bf16_t st_via_sr(C v)
    {
        const float f = float(v);
        const uint32_t w = (0x7fffu + ((__float_as_uint(f) >> 16) & 1u)) << 16;
        return bf16_t{uint16_t(__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(bf16x2_t{}, f, w, false)))};
    }
};

// 32-bit integer mix (C. Wellons' published "lowbias32")
__device__ __forceinline__ uint32_t mix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x7feb352du;
    h ^= h >> 15;
    h *= 0x846ca68bu;
    h ^= h >> 16;
    return h;
}

// Stochastic rounding to BF16: the FP32 value's low 16 bits are added to
// the dither (the high half of word) and the sum is truncated, so it rounds
// away from zero with probability equal to the dropped fraction and the
// rounding adds no bias (a mean error of zero, against round-to-nearest's
// drift on the slowly varying populations).  One instruction, gfx950's
// v_cvt_sr_bf16_f32, which is this sum for every finite input (checked
// against the bit formula on 10^6 inputs); a finite value carried past the
// largest BF16 becomes infinity, NaN stays NaN.
template <>
struct Store<bf16sr_t> {
    __device__ static float ld(bf16sr_t v) { return __uint_as_float(uint32_t(v.bits) << 16); }
    template <class C>
    __device__ static bf16sr_t st(C v, uint32_t word = 0u)
    {
        return bf16sr_t{uint16_t(__builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(bf16x2_t{}, float(v), word, false)))};
    }
};

// The rounding words of a cell's Q stored values, sr_word(...) + q * kGolden:
// one hash per cell and step of the cell's index and its pressure (which
// changes every step), then a golden-ratio (Weyl) sequence over the
// directions, so each value's dither is uniform and the Q dithers of a cell
// are evenly spread.  Deterministic: runs repeat bit for bit.  Zero (and no
// work) for every other storage form.
constexpr uint32_t kGolden = 0x9e3779b9u;
template <class S, class C>
__device__ __forceinline__ uint32_t sr_word(C p, uint32_t c)
{
    if constexpr (std::is_same<S, bf16sr_t>::value)
        return mix32(__float_as_uint(float(p)) ^ (c * kGolden));
    else
        return 0u;
}

// streaming (non-temporal) store of one population value
template <class S>
__device__ __forceinline__ void nt_store(S *p, S v)
{
    __builtin_nontemporal_store(v, p);
}
template <>
__device__ __forceinline__ void nt_store<bf16_t>(bf16_t *p, bf16_t v)
{
    __builtin_nontemporal_store(v.bits, &p->bits);
}
template <>
__device__ __forceinline__ void nt_store<bf16sr_t>(bf16sr_t *p, bf16sr_t v)
{
    __builtin_nontemporal_store(v.bits, &p->bits);
}
template <class S>
__device__ __forceinline__ S nt_load(const S *p)
{
    return __builtin_nontemporal_load(p);
}
template <>
__device__ __forceinline__ bf16_t nt_load<bf16_t>(const bf16_t *p)
{
    return bf16_t{__builtin_nontemporal_load(&p->bits)};
}
template <>
__device__ __forceinline__ bf16sr_t nt_load<bf16sr_t>(const bf16sr_t *p)
{
    return bf16sr_t{__builtin_nontemporal_load(&p->bits)};
}

// ---------------------------------------------------------------------------
// node decoding
__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &d)
{
    return uint32_t((uint64_t(__umulhi(n, d.m)) + n) >> d.s);
}

// Blocks are dealt round-robin to the 8 XCDs, so two neighbouring blocks
// run on different XCDs and each fetches the cache lines they share into
// its own L2.  remap = 1 gives each XCD one contiguous eighth of the
// blocks; remap = R >= 2 gives each XCD runs of R consecutive blocks, the
// eight XCDs working on one window of 8R blocks at a time; 0 keeps the
// hardware order.  The lattice context chooses (measured on MI355X).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb, uint32_t remap)
{
    if (remap == 0) return b;
    if (remap == 1) {
        const uint32_t per = nb / 8u;
        if (b >= per * 8u) return b;
        return (b % 8u) * per + b / 8u;
    }
    const uint32_t span = 8u * remap;
    if (b >= nb / span * span) return b;  // the last partial window in hardware order
    const uint32_t r = b % span;          // XCD r % 8, position r / 8 of its run
    return b - r + (r % 8u) * remap + r / 8u;
}

template <int D>
__device__ __forceinline__ bool decode(const KParams &kp, uint32_t i, int &x, int &y, int &z, uint32_t &c)
{
    if (i >= kp.N) return false;
    const uint32_t r = fdiv(i, kp.div_x);
    x = int(i - r * uint32_t(kp.nx + 1));
    if (D == 3) {
        const uint32_t zz = fdiv(r, kp.div_y);
        y = int(r - zz * uint32_t(kp.ny + 1));
        z = int(zz);
    } else {
        y = int(r);
        z = 0;
    }
    c = uint32_t(kp.o0 + int64_t(x) + kp.Px * y + kp.PxPy * z);
    return true;
}

// Cell-mapped threads: thread i is padded cell c = i; threads on ghost or
// padding cells leave.  Waves then cover whole tiles of the tiled layout.
template <int D>
__device__ __forceinline__ bool decode_cell(const KParams &kp, uint32_t i, int &x, int &y, int &z, uint32_t &c)
{
    if (i >= kp.ncell) return false;
    const uint32_t r = fdiv(i, kp.div_px);
    x = int(i - r * uint32_t(kp.Px)) - 1;
    if (D == 3) {
        const uint32_t zz = fdiv(r, kp.div_py);
        y = int(r - zz * uint32_t(kp.Py)) - 1;
        z = int(zz) - 1;
    } else {
        y = int(r) - 1;
        z = 0;
    }
    c = i;
    return x >= 0 && x <= kp.nx && y >= 0 && y <= kp.ny && (D == 2 || (z >= 0 && z <= kp.nz));
}

// node of thread i in the configured thread mapping
template <int D>
__device__ __forceinline__ bool node_of(const KParams &kp, uint32_t i, int &x, int &y, int &z, uint32_t &c)
{
    return IMEXLB_CELLMAP ? decode_cell<D>(kp, i, x, y, z, c) : decode<D>(kp, i, x, y, z, c);
}

template <int D>
__device__ __forceinline__ int64_t shift(const KParams &kp, int q)
{
    using L = Lat<D>;
    return int64_t(L::ex(q)) + kp.Px * L::ey(q) + kp.PxPy * L::ez(q);
}

// Population slot of direction q at padded cell c: planes (one per q) or
// tiles of IMEXLB_TILE cells holding all directions of the tile together.
// Unsigned cell index: no sign fix-ups in the tile division, and the slot
// comes out of one 64-bit multiply-add.
template <int Q>
__device__ __forceinline__ uint64_t cslot(const KParams &kp, uint32_t q, uint32_t cell)
{
#if IMEXLB_TILE
    constexpr uint32_t T = IMEXLB_TILE;
    return uint64_t(cell / T) * (uint64_t(Q) * T) + ((cell % T) | (uint32_t(store_q(int(q), Q)) * T));
#else
    return uint64_t(q) * uint64_t(kp.plane) + cell;
#endif
}

// padded-cell offset of the neighbour a population of direction q comes
// from, e_q in cells (|offset| < 2^31)
template <int D>
__device__ __forceinline__ uint32_t shift32(const KParams &kp, int q)
{
    using L = Lat<D>;
    return uint32_t(L::ex(q) + int32_t(kp.Px) * L::ey(q) + int32_t(kp.PxPy) * L::ez(q));
}

// ---------------------------------------------------------------------------
// FP32-compute forms (FP16/BF16/FP32 storage): the same quantities with
// fewer instructions -- no bit-for-bit constraint applies, so sums start at
// their first term, divisions become multiplies by reciprocals and products
// fuse into FMAs.

// sum of f[q] over the directions where pick(q) holds (0 if none)
template <int Q, class Pick>
__device__ __forceinline__ float sum_where(const float (&f)[Q], Pick pick)
{
    float acc = 0.0f;
    bool any = false;
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (pick(q)) {
            acc = any ? acc + f[q] : f[q];
            any = true;
        }
    return acc;
}

template <int D>
__device__ __forceinline__ void moments_f32(const float (&f)[Lat<D>::Q], const KParams &kp, int x, int y, int z,
                                            float &p, float &ux, float &uy, float &uz)
{
    using L = Lat<D>;
    const float ps = sum_where<L::Q>(f, [](int) { return true; });
    const float sx = sum_where<L::Q>(f, [](int q) { return L::ex(q) > 0; }) -
                     sum_where<L::Q>(f, [](int q) { return L::ex(q) < 0; });
    const float sy = sum_where<L::Q>(f, [](int q) { return L::ey(q) > 0; }) -
                     sum_where<L::Q>(f, [](int q) { return L::ey(q) < 0; });
    const float ir = kp.inv_rho0_f;
    p = ps * (1.0f / 3.0f);
    ux = sx * ir;
    uy = sy * ir;
    uz = 0.0f;
    if (D == 3)
        uz = (sum_where<L::Q>(f, [](int q) { return L::ez(q) > 0; }) -
              sum_where<L::Q>(f, [](int q) { return L::ez(q) < 0; })) *
             ir;
    if (x == 0) {
        const int h = D == 2 ? y : z;
        const int nh = D == 2 ? kp.ny : kp.nz;
        ux = kp.uu_f * kp.rho0_f * 4.0f * float(h) * float(nh - h) / float(nh * nh);
        uy = 0.0f;
    }
    if (x == kp.nx) p = 0.0f;
}

// feq_q = w_q (3p + rho0 (3 e.u + 4.5 (e.u)^2 - 1.5 u.u)) and
// f' = (1-ome) f + ome feq as (1-ome) f + (ome w_q rho0) t + ome w_q base with
// t = e.u (3 + 4.5 e.u) and base = 3p - 1.5 rho0 u.u; the per-weight
// products are shared between directions of equal weight.
template <int D, class Emit>
__device__ __forceinline__ void collide_f32(const float (&f)[Lat<D>::Q], float p, float ux, float uy, float uz,
                                            const KParams &kp, Emit emit)
{
    using L = Lat<D>;
    const float ome = kp.ome_f, rho0 = kp.rho0_f;
    const float omc = 1.0f - ome;
    float udu = ux * ux;
    udu = __builtin_fmaf(uy, uy, udu);
    if (D == 3) udu = __builtin_fmaf(uz, uz, udu);
    const float base = __builtin_fmaf(-1.5f * rho0, udu, 3.0f * p);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        const float ow = ome * float(L::w(q));
        float edu = 0.0f;
        bool first = true;
        auto term = [&](int e, float v) {
            if (e == 0) return;
            edu = first ? (e > 0 ? v : -v) : (e > 0 ? edu + v : edu - v);
            first = false;
        };
        term(L::ex(q), ux);
        term(L::ey(q), uy);
        if (D == 3) term(L::ez(q), uz);
        float eq = ow * base;
        if (!first) eq = __builtin_fmaf(ow * rho0, edu * __builtin_fmaf(4.5f, edu, 3.0f), eq);
        emit(q, __builtin_fmaf(omc, f[q], eq));
    }
}

// ---------------------------------------------------------------------------
// PostProcessing arithmetic on one node (lbm.f90:163-199): moments summed in
// q order, divided by 3 and by rho0, then the inlet Dirichlet velocity at
// x = 0 and the outlet p = 0 at x = nx.
template <int D, class C>
__device__ __forceinline__ void moments(const C (&f)[Lat<D>::Q], const KParams &kp, int x, int y, int z, C &p,
                                        C &ux, C &uy, C &uz)
{
    using L = Lat<D>;
    if constexpr (std::is_same<C, float>::value) {
        moments_f32<D>(f, kp, x, y, z, p, ux, uy, uz);
        return;
    }
    C ps = C(0), sx = C(0), sy = C(0), sz = C(0);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        ps = ps + f[q];
        if (L::ex(q) > 0) sx = sx + f[q];
        if (L::ex(q) < 0) sx = sx - f[q];
        if (L::ey(q) > 0) sy = sy + f[q];
        if (L::ey(q) < 0) sy = sy - f[q];
        if (D == 3 && L::ez(q) > 0) sz = sz + f[q];
        if (D == 3 && L::ez(q) < 0) sz = sz - f[q];
    }
    const C rho0 = C(kp.rho0);
    p = ps / C(3);
    ux = sx / rho0;
    uy = sy / rho0;
    uz = D == 3 ? sz / rho0 : C(0);
    if (x == 0) {  // inlet, lbm.f90:186-191 (3D: profile in z, u_z kept, lbm_3d.f90:188-193)
        const int h = D == 2 ? y : z;
        const int nh = D == 2 ? kp.ny : kp.nz;
        ux = C(kp.uu) * rho0 * C(4) * C(h) * C(nh - h) / C(nh * nh);
        uy = C(0);
    }
    if (x == kp.nx) p = C(0);  // outlet, lbm.f90:195-198
}

// Equilibrium (lbm.f90:70) and BGK relaxation (lbm.f90:71) of one node.
// e.u is formed term by term in x, y, z order with the zero terms of the
// reference's sum dropped (they only change the sign of a zero).
template <int D, class C, class Emit>
__device__ __forceinline__ void collide(const C (&f)[Lat<D>::Q], C p, C ux, C uy, C uz, const KParams &kp,
                                        Emit emit)
{
    using L = Lat<D>;
    if constexpr (std::is_same<C, float>::value) {
        collide_f32<D>(f, p, ux, uy, uz, kp, emit);
        return;
    }
    const C ome = C(kp.ome), rho0 = C(kp.rho0);
    const C omc = C(1) - ome;
    C udu = ux * ux + uy * uy;
    if (D == 3) udu = udu + uz * uz;
    const C p3 = p * C(3);
    const C u15 = C(1.5) * udu;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        C edu = C(0);
        bool first = true;
        auto term = [&](int e, C v) {
            if (e == 0) return;
            if (first) {
                edu = e > 0 ? v : -v;
                first = false;
            } else {
                edu = e > 0 ? edu + v : edu - v;
            }
        };
        term(L::ex(q), ux);
        term(L::ey(q), uy);
        if (D == 3) term(L::ez(q), uz);
        const C feq = C(L::w(q)) * (p3 + rho0 * ((C(3) * edu + C(4.5) * (edu * edu)) - u15));
        emit(q, omc * f[q] + ome * feq);
    }
}

// ---------------------------------------------------------------------------
// Order in which the gather issues its loads: q order (0), or grouped by the
// source row (y, z shift), x shift fastest (1), so consecutive loads of a
// wave fall in the same stretch of memory.  Arithmetic keeps q order.
#ifndef IMEXLB_LOAD_ORDER
#define IMEXLB_LOAD_ORDER 0
#endif
template <int D>
__device__ __forceinline__ constexpr int load_order(int j)
{
    if (!IMEXLB_LOAD_ORDER) return j;
    return D == 2 ? (j % 3) * 3 + j / 3 : (j / 9) * 9 + (j % 3) * 3 + (j % 9) / 3;
}

// ---------------------------------------------------------------------------
// The step kernel.  MODE bits:
//   K_COLLIDE  relax and write dst (the next step's post-collision state)
//   K_WRITE_F  write the gathered populations (array f)
//   K_WRITE_PU write p and u (moments after the inlet/outlet conditions)
//   K_NOSHIFT  read src at the node itself (PostProcessing of f in place)
// COLLIDE alone is the fused Propagation+PostProcessing+Collision pull;
// WRITE_F alone is Propagation; NOSHIFT|WRITE_PU is PostProcessing;
// WRITE_F|WRITE_PU materialises f, p, u lazily.
template <int MODE>
__device__ __forceinline__ uint64_t push_slot(const PushTable &pt, int32_t e)
{
    return (MODE & K_IDX64) ? uint64_t(static_cast<const int64_t *>(pt.dst)[e])
                            : uint64_t(static_cast<const int32_t *>(pt.dst)[e]);
}

// The gather half of a node update: the Q populations the node pulls (or
// its own with K_NOSHIFT).  Every gathered slot exists in the arrays also
// for solid nodes, so the loads need not wait for the node's type.
template <int D, class S, class C, int MODE>
__device__ __forceinline__ void gather(const S *__restrict__ src, const KParams &kp, uint32_t c,
                                       C (&f)[Lat<D>::Q])
{
    using L = Lat<D>;
#pragma unroll
    for (int j = 0; j < L::Q; ++j) {
        const int q = load_order<D>(j);
        const S *a = src + cslot<L::Q>(kp, q, (MODE & K_NOSHIFT) ? c : c - shift32<D>(kp, q));
        f[q] = Store<S>::ld(IMEXLB_STEP_NTLOAD ? nt_load(a) : *a);
    }
}

// Internal mode bit of the step kernel: the wave covers exactly one 64-cell
// tile (cell-mapped threads, 64-cell tiles, one cell per thread), so slots
// are formed from a wave-uniform tile base plus a 32-bit lane offset.
constexpr int K_WAVE = 64;
#ifndef IMEXLB_ADDR32
#define IMEXLB_ADDR32 0
#endif
constexpr bool kWaveSlots = IMEXLB_ADDR32 && IMEXLB_CELLMAP && IMEXLB_TILE == 64;

// the tile of the wave that holds cell c (every lane of the wave has it)
__device__ __forceinline__ uint32_t wave_tile(uint32_t c) { return __builtin_amdgcn_readfirstlane(c >> 6); }

// gather() for K_WAVE: the populations the node pulls, addressed from the
// tile back_tiles behind the wave's own, the lowest any pull can reach, so
// every offset is a non-negative 32-bit count of values.
template <int D, class S, class C>
__device__ __forceinline__ void gather_wave(const S *__restrict__ src, const KParams &kp, uint32_t c,
                                            C (&f)[Lat<D>::Q])
{
    using L = Lat<D>;
    constexpr uint32_t QT = uint32_t(L::Q) * 64u;
    const uint32_t back = kp.back_tiles;
    // near the start of the array the base lies before it; only base + off
    // is ever read (pointer arithmetic keeps the global address space)
    const S *base = src + (int64_t(wave_tile(c)) - int64_t(back)) * int64_t(QT);
    const uint32_t lane = c & 63u;
#pragma unroll
    for (int j = 0; j < L::Q; ++j) {
        const int q = load_order<D>(j);
        // source cell relative to the start of the base tile; byte offsets
        // in 32 bits (tile count < 2^24 and the product < 2^32 for any
        // lattice whose slots fit 32-bit cells), the direction's position
        // in the tile as a constant added last
        const uint32_t rel = lane + back * 64u - shift32<D>(kp, q);
        const uint32_t boff = __umul24(rel >> 6, QT * uint32_t(sizeof(S))) + (rel & 63u) * uint32_t(sizeof(S));
        const S *a = reinterpret_cast<const S *>(reinterpret_cast<const char *>(base) + boff) +
                     uint32_t(store_q(q, L::Q)) * 64u;
        f[q] = Store<S>::ld(IMEXLB_STEP_NTLOAD ? nt_load(a) : *a);
    }
}

// Entries [b, e) of the push table that cell c copies out: empty unless
// K_PUSH and the cell has copies.
struct PushRange {
    int32_t b, e;
};
template <int MODE>
__device__ __forceinline__ PushRange push_range(const PushTable &pt, uint8_t g, uint32_t c)
{
    PushRange r{0, 0};
    if ((MODE & K_PUSH) && (g & G_PUSH)) {
        r.b = pt.beg[c];
        r.e = pt.beg[c + 1];
    }
    return r;
}

// v[q] for a direction known only at run time, as an OR of masked bit
// patterns: no indexed access, which would move v to scratch memory
template <int Q, class S>
__device__ __forceinline__ S pick(const S (&v)[Q], uint32_t q)
{
    using U = std::conditional_t<sizeof(S) == 8, uint64_t, std::conditional_t<sizeof(S) == 4, uint32_t, uint16_t>>;
    U r = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k) r |= __builtin_bit_cast(U, v[k]) & U(U(0) - U(q == uint32_t(k)));
    return __builtin_bit_cast(S, r);
}

// Boundary copies of D2Q9 nodes are written from registers (the node's
// output values, or a body cell's own values loaded at once) with the first
// IMEXLB_PUSH_PREFETCH table entries loaded before the node's arithmetic:
// on small lattices the copies are the longest chain of dependent memory
// accesses in the step.  D3Q27 keeps one read of the value per copy.
#ifndef IMEXLB_PUSH_PREFETCH
#define IMEXLB_PUSH_PREFETCH 4
#endif

// The rest of a node update from its gathered populations and node type g.
template <int D, class S, class C, int MODE>
__device__ __forceinline__ void finish_node(const C (&f)[Lat<D>::Q], uint8_t g, const S *__restrict__ src,
                                            S *__restrict__ dst, S *__restrict__ fout, double *__restrict__ pout,
                                            double *__restrict__ uxo, double *__restrict__ uyo,
                                            double *__restrict__ uzo, const KParams &kp, int x, int y, int z,
                                            uint32_t c, const PushTable &pt);

template <int D, class S, class C, int MODE>
__device__ __forceinline__ void step_node(const S *__restrict__ src, S *__restrict__ dst, S *__restrict__ fout,
                                          double *__restrict__ pout, double *__restrict__ uxo,
                                          double *__restrict__ uyo, double *__restrict__ uzo,
                                          const uint8_t *__restrict__ geo, const KParams &kp, int x, int y, int z,
                                          uint32_t c, const PushTable &pt = PushTable{})
{
    // the node's type is loaded together with its populations, not before
    const uint8_t g = geo[c];
    C f[Lat<D>::Q];
    gather<D, S, C, MODE>(src, kp, c, f);
    finish_node<D, S, C, MODE>(f, g, src, dst, fout, pout, uxo, uyo, uzo, kp, x, y, z, c, pt);
}

template <int D, class S, class C, int MODE>
__device__ __forceinline__ void finish_node(const C (&f)[Lat<D>::Q], uint8_t g, const S *__restrict__ src,
                                            S *__restrict__ dst, S *__restrict__ fout, double *__restrict__ pout,
                                            double *__restrict__ uxo, double *__restrict__ uyo,
                                            double *__restrict__ uzo, const KParams &kp, int x, int y, int z,
                                            uint32_t c, const PushTable &pt)
{
    using L = Lat<D>;
    const bool fluid = (g & G_FLUID) != 0;
    // D2Q9 copies: the cell's range and its first entries, before any arithmetic
    constexpr bool kRegPush = (MODE & K_PUSH) && D == 2;
    constexpr int PD = kRegPush ? IMEXLB_PUSH_PREFETCH : 0;
    PushRange pr{0, 0};
    uint32_t pq[PD > 0 ? PD : 1];
    uint64_t ps[PD > 0 ? PD : 1];
    if constexpr (kRegPush) {
        pr = push_range<MODE>(pt, g, c);
#pragma unroll
        for (int k = 0; k < PD; ++k)
            if (pr.b + k < pr.e) {
                pq[k] = pt.q[pr.b + k];
                ps[k] = push_slot<MODE>(pt, pr.b + k);
            }
    }
    if (!fluid) {
        if (MODE & K_WRITE_PU) {  // PostProcessing's inlet/outlet loops also reach solid nodes
            if (x == 0) {
                const int h = D == 2 ? y : z;
                const int nh = D == 2 ? kp.ny : kp.nz;
                uxo[c] = double(C(kp.uu) * C(kp.rho0) * C(4) * C(h) * C(nh - h) / C(nh * nh));
                uyo[c] = 0.0;
            }
            if (x == kp.nx) pout[c] = 0.0;
        }
        if constexpr (kRegPush) {  // a body cell: its previous-step values
            if (pr.b < pr.e) {
                S own[L::Q];
#pragma unroll
                for (int q = 0; q < L::Q; ++q) own[q] = src[cslot<L::Q>(kp, q, c)];
#pragma unroll
                for (int k = 0; k < PD; ++k)
                    if (pr.b + k < pr.e) dst[ps[k]] = pick(own, pq[k]);
                for (int32_t e = pr.b + PD; e < pr.e; ++e) dst[push_slot<MODE>(pt, e)] = pick(own, pt.q[e]);
            }
        } else if ((MODE & K_PUSH) && (g & G_PUSH)) {
            const int32_t e1 = pt.beg[c + 1];
            for (int32_t e = pt.beg[c]; e < e1; ++e) dst[push_slot<MODE>(pt, e)] = src[cslot<L::Q>(kp, pt.q[e], c)];
        }
        return;
    }
    if (MODE & K_WRITE_F) {
#pragma unroll
        for (int q = 0; q < L::Q; ++q) fout[cslot<L::Q>(kp, q, c)] = Store<S>::template st<C>(f[q]);
    }
    if (!(MODE & (K_WRITE_PU | K_COLLIDE))) return;
    C p, ux, uy, uz;
    moments<D, C>(f, kp, x, y, z, p, ux, uy, uz);
    if (MODE & K_WRITE_PU) {
        pout[c] = double(p);
        uxo[c] = double(ux);
        uyo[c] = double(uy);
        if (D == 3) uzo[c] = double(uz);
    }
    if (MODE & K_COLLIDE) {
        // K_WAVE: the wave is one tile; its output slots are a scalar tile
        // base plus the lane
        S *const wdst = (MODE & K_WAVE) ? dst + uint64_t(wave_tile(c)) * (L::Q * 64u) : dst;
        S out[kRegPush ? L::Q : 1];  // the stored values, for the copies
        const uint32_t srw = sr_word<S>(p, c);
        collide<D, C>(f, p, ux, uy, uz, kp, [&](int q, C v) {
            S *o = (MODE & K_WAVE) ? wdst + (uint32_t(store_q(q, L::Q)) * 64u + (c & 63u)) : dst + cslot<L::Q>(kp, q, c);
            const S sv = Store<S>::template st<C>(v, srw + uint32_t(q) * kGolden);
            if constexpr (kRegPush) out[q] = sv;
            if (IMEXLB_STEP_NT)
                nt_store(o, sv);
            else
                *o = sv;
        });
        if constexpr (kRegPush) {  // boundary copies of this node's output
            if (pr.b < pr.e) {
#pragma unroll
                for (int k = 0; k < PD; ++k)
                    if (pr.b + k < pr.e) dst[ps[k]] = pick(out, pq[k]);
                for (int32_t e = pr.b + PD; e < pr.e; ++e) dst[push_slot<MODE>(pt, e)] = pick(out, pt.q[e]);
            }
        } else if ((MODE & K_PUSH) && (g & G_PUSH)) {  // boundary copies of this node's output (read back)
            const int32_t e1 = pt.beg[c + 1];
            for (int32_t e = pt.beg[c]; e < e1; ++e) dst[push_slot<MODE>(pt, e)] = dst[cslot<L::Q>(kp, pt.q[e], c)];
        }
    }
}

// K_SMALL step (D2Q9 lattices small enough to stay in the caches between
// steps, where the time of a step is the time of its slowest wave): a block
// is one wave and one 64-cell tile.  The push table is grouped by cell, so
// the tile's boundary copies are one run [eb, ee) of it, found with two
// loads at wave-uniform addresses that need nothing from the cells.  Every
// lane of a tile with copies leaves its values in LDS (a fluid node its
// output, a body cell its previous-step values), and the run is spread over
// the wave's lanes: one LDS read and one store per copy, the first
// kSmallRounds rounds of table entries (and in tiles with copies every
// cell's own values) loaded before the arithmetic.  Taking the entries into
// registers before the node's output stores, or storing the outputs after
// the arithmetic, measured slower.
constexpr int kSmallRounds = 4;
template <class S>
using BitsOf = std::conditional_t<sizeof(S) == 8, uint64_t, std::conditional_t<sizeof(S) == 4, uint32_t, uint16_t>>;

template <int D, class S, class C, int MODE>
__device__ __forceinline__ void step_small(const S *__restrict__ src, S *__restrict__ dst,
                                           const uint8_t *__restrict__ geo, const KParams &kp, const PushTable &pt,
                                           uint32_t tile)
{
    static_assert(D == 2 && (MODE & K_PUSH) && !(MODE & K_IDX64), "K_SMALL: D2Q9 with int32 push slots");
    static_assert(IMEXLB_CELLMAP && kSmallBlock == 64 && IMEXLB_TILE == 64,
                  "K_SMALL: one thread per padded cell, one-wave blocks of one 64-cell tile");
    using L = Lat<D>;
    using U = BitsOf<S>;
    __shared__ U sh[L::Q * 64];
    const uint32_t lane = threadIdx.x;
    const uint32_t c0 = tile * 64u;
    if (c0 >= kp.ncell) return;  // the whole block
    const int32_t eb = __builtin_amdgcn_readfirstlane(pt.beg[c0]);
    const int32_t ee = __builtin_amdgcn_readfirstlane(pt.beg[min(c0 + 64u, kp.ncell)]);
    const int32_t *const pdst = static_cast<const int32_t *>(pt.dst);
    uint32_t li[kSmallRounds] = {}, ds[kSmallRounds] = {};
#pragma unroll
    for (int r = 0; r < kSmallRounds; ++r) {
        const int32_t e = eb + int32_t(lane) + 64 * r;
        if (e < ee) {
            li[r] = pt.lds[e];
            ds[r] = uint32_t(pdst[e]);
        }
    }
    // the tile's slots from a wave-uniform base: slot q of lane l is
    // base[store_q(q) * 64 + l] (cslot of cell c0 + l)
    const S *const tsrc = src + uint64_t(tile) * (L::Q * 64u);
    S *const tdst = dst + uint64_t(tile) * (L::Q * 64u);
    int x, y, z;
    uint32_t c;
    if (decode_cell<D>(kp, c0 + lane, x, y, z, c)) {
        const uint8_t g = geo[c];
        C f[L::Q];
        gather<D, S, C, MODE>(src, kp, c, f);
        U own[L::Q];
        if (eb < ee) {  // what a body cell copies, loaded with the populations
#pragma unroll
            for (int q = 0; q < L::Q; ++q) own[q] = __builtin_bit_cast(U, tsrc[uint32_t(store_q(q, L::Q)) * 64u + lane]);
        }
        if (g & G_FLUID) {
            C p, ux, uy, uz;
            moments<D, C>(f, kp, x, y, z, p, ux, uy, uz);
            const uint32_t srw = sr_word<S>(p, c);
            collide<D, C>(f, p, ux, uy, uz, kp, [&](int q, C v) {
                const S sv = Store<S>::template st<C>(v, srw + uint32_t(q) * kGolden);
                tdst[uint32_t(store_q(q, L::Q)) * 64u + lane] = sv;
                if (eb < ee) sh[q * 64 + int(lane)] = __builtin_bit_cast(U, sv);
            });
        } else if (g & G_PUSH) {  // (a cell with copies: eb < ee)
#pragma unroll
            for (int q = 0; q < L::Q; ++q) sh[q * 64 + int(lane)] = own[q];
        }
    }
    if (eb < ee) {  // the same for every lane of the block
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSmallRounds; ++r)
            if (eb + int32_t(lane) + 64 * r < ee) dst[ds[r]] = __builtin_bit_cast(S, sh[li[r]]);
        for (int32_t e = eb + int32_t(lane) + 64 * kSmallRounds; e < ee; e += 64)
            dst[uint32_t(pdst[e])] = __builtin_bit_cast(S, sh[pt.lds[e]]);
    }
}

// The fused step of a D3Q27 lattice in 16- or 32-bit storage, one wave per
// 64-cell tile [c0, c0 + 64): as step_small, each wave with boundary copies
// leaves its cells' values in LDS (27 x 64 values) and spreads the tile's
// run of the push table over its lanes, one LDS read and one store per copy,
// instead of each cell reading back its own output per copy.  Output stores
// non-temporal.  (FP64 keeps the read-back: 13.8 KB of LDS per wave would
// cost the kernel a wave per SIMD.)
template <int D, class S, class C, int MODE>
__device__ __forceinline__ void step_wave_copies(const S *__restrict__ src, S *__restrict__ dst,
                                                 const uint8_t *__restrict__ geo, const KParams &kp,
                                                 const PushTable &pt, uint32_t c0, uint32_t lane, BitsOf<S> *sh)
{
    static_assert(IMEXLB_CELLMAP && IMEXLB_TILE == 64 && (MODE & K_PUSH) && !(MODE & K_IDX64),
                  "one wave per 64-cell tile, int32 push slots");
    using L = Lat<D>;
    using U = BitsOf<S>;
    if (c0 >= kp.ncell) return;  // the whole wave
    const int32_t eb = __builtin_amdgcn_readfirstlane(pt.beg[c0]);
    const int32_t ee = __builtin_amdgcn_readfirstlane(pt.beg[min(c0 + 64u, kp.ncell)]);
    const int32_t *const pdst = static_cast<const int32_t *>(pt.dst);
    uint32_t li[kSmallRounds] = {}, ds[kSmallRounds] = {};
#pragma unroll
    for (int r = 0; r < kSmallRounds; ++r) {
        const int32_t e = eb + int32_t(lane) + 64 * r;
        if (e < ee) {
            li[r] = pt.lds[e];
            ds[r] = uint32_t(pdst[e]);
        }
    }
    // the tile's slots from a wave-uniform base: slot q of lane l is
    // base[store_q(q) * 64 + l] (cslot of cell c0 + l)
    const uint32_t tile = __builtin_amdgcn_readfirstlane(c0 >> 6);
    const S *const tsrc = src + uint64_t(tile) * (L::Q * 64u);
    S *const tdst = dst + uint64_t(tile) * (L::Q * 64u);
    int x, y, z;
    uint32_t c;
    if (decode_cell<D>(kp, c0 + lane, x, y, z, c)) {
        const uint8_t g = geo[c];
        C f[L::Q];
        gather<D, S, C, MODE>(src, kp, c, f);
        if (g & G_FLUID) {
            C p, ux, uy, uz;
            moments<D, C>(f, kp, x, y, z, p, ux, uy, uz);
            const uint32_t srw = sr_word<S>(p, c);
            collide<D, C>(f, p, ux, uy, uz, kp, [&](int q, C v) {
                S sv;
                if constexpr (std::is_same<S, bf16_t>::value)
                    sv = Store<S>::template st_via_sr<C>(v);
                else
                    sv = Store<S>::template st<C>(v, srw + uint32_t(q) * kGolden);
                S *const o = tdst + (uint32_t(store_q(q, L::Q)) * 64u + lane);
                if (IMEXLB_STEP_NT)
                    nt_store(o, sv);
                else
                    *o = sv;
                if (eb < ee) sh[q * 64 + int(lane)] = __builtin_bit_cast(U, sv);
            });
        } else if (g & G_PUSH) {  // a body cell: its previous-step values
#pragma unroll
            for (int q = 0; q < L::Q; ++q)
                sh[q * 64 + int(lane)] = __builtin_bit_cast(U, tsrc[uint32_t(store_q(q, L::Q)) * 64u + lane]);
        }
    }
    if (eb < ee) {  // the same for every lane of the wave
        // the copied values are other lanes' LDS writes of this wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int r = 0; r < kSmallRounds; ++r)
            if (eb + int32_t(lane) + 64 * r < ee) dst[ds[r]] = __builtin_bit_cast(S, sh[li[r]]);
        for (int32_t e = eb + int32_t(lane) + 64 * kSmallRounds; e < ee; e += 64)
            dst[uint32_t(pdst[e])] = __builtin_bit_cast(S, sh[pt.lds[e]]);
    }
}

// Threads per block of the fused step: one wave for the small-lattice form
// and for D3Q27 in FP32 storage (2.8 % faster on the sphere than 256), 128
// for D2Q9 in FP64 storage (2.5 % faster on the 8192x2048 cylinder), and
// IMEXLB_STEP_BLOCK otherwise (the 16-bit forms lose 2-4 % at 128 and
// 11-18 % at 64, FP64 D3Q27 0.4-0.6 % at 64 and 128).
template <int D, class S, int MODE>
__host__ __device__ constexpr unsigned step_block()
{
    return (MODE & K_SMALL) || (D == 3 && sizeof(S) == 4) ? unsigned(kSmallBlock)
           : (D == 2 && sizeof(S) == 8 && IMEXLB_STEP_BLOCK >= 128) ? 128u
                                                                   : unsigned(IMEXLB_STEP_BLOCK);
}

#if IMEXLB_STEP_WAVES
#define IMEXLB_STEP_BOUNDS __launch_bounds__((step_block<D, S, MODE>()), IMEXLB_STEP_WAVES)
#else
#define IMEXLB_STEP_BOUNDS __launch_bounds__((step_block<D, S, MODE>()))
#endif

// Cells per thread of the fused collide step: a wave takes CPT consecutive
// 64-cell tiles, lane l cell l of each, and issues the loads of all of them
// before any arithmetic (more bytes in flight per wave).
template <int MODE>
__host__ __device__ constexpr int step_cpt()
{
    return (MODE & K_COLLIDE) && !(MODE & (K_WRITE_F | K_WRITE_PU | K_SMALL)) ? IMEXLB_STEP_CPT : 1;
}

template <int D, class S, class C, int MODE>
__global__ IMEXLB_STEP_BOUNDS void k_step(const S *__restrict__ src, S *__restrict__ dst,
                                              S *__restrict__ fout, double *__restrict__ pout,
                                              double *__restrict__ uxo, double *__restrict__ uyo,
                                              double *__restrict__ uzo, const uint8_t *__restrict__ geo,
                                              KParams kp, PushTable pt)
{
    constexpr int CPT = step_cpt<MODE>();
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x, kp.xcd_remap);
    if constexpr ((MODE & K_SMALL) != 0) {
        step_small<D, S, C, MODE>(src, dst, geo, kp, pt, blk);
    } else if constexpr (D == 3 && sizeof(S) <= 4 && MODE == (K_COLLIDE | K_PUSH) && CPT == 1 && IMEXLB_CELLMAP &&
                         IMEXLB_TILE == 64) {
        __shared__ BitsOf<S> sh[step_block<D, S, MODE>() / 64][Lat<D>::Q * 64];
        step_wave_copies<D, S, C, MODE>(src, dst, geo, kp, pt, blk * step_block<D, S, MODE>() + (threadIdx.x & ~63u),
                                        threadIdx.x & 63u, sh[threadIdx.x >> 6]);
    } else if constexpr (CPT == 1) {
        int x, y, z;
        uint32_t c;
        const uint32_t i = blk * blockDim.x + threadIdx.x;
        if (!node_of<D>(kp, i, x, y, z, c)) return;
        if constexpr (kWaveSlots && MODE == (MODE & (K_COLLIDE | K_PUSH | K_IDX64))) {
            const uint8_t g = geo[c];
            C f[Lat<D>::Q];
            gather_wave<D, S, C>(src, kp, c, f);
            finish_node<D, S, C, MODE | K_WAVE>(f, g, src, dst, fout, pout, uxo, uyo, uzo, kp, x, y, z, c, pt);
        } else if constexpr (IMEXLB_CELLMAP && IMEXLB_TILE == 64 && MODE == (MODE & (K_COLLIDE | K_PUSH | K_IDX64))) {
            // the wave is one tile: output slots from its wave-uniform base
            const uint8_t g = geo[c];
            C f[Lat<D>::Q];
            gather<D, S, C, MODE>(src, kp, c, f);
            finish_node<D, S, C, MODE | K_WAVE>(f, g, src, dst, fout, pout, uxo, uyo, uzo, kp, x, y, z, c, pt);
        } else {
            step_node<D, S, C, MODE>(src, dst, fout, pout, uxo, uyo, uzo, geo, kp, x, y, z, c, pt);
        }
    } else {
        const uint32_t lane = threadIdx.x & 63u;
        const uint32_t wave = (blk * blockDim.x + threadIdx.x) >> 6;
        int x[CPT], y[CPT], z[CPT];
        uint32_t c[CPT];
        bool ok[CPT];
        uint8_t g[CPT];
        C f[CPT][Lat<D>::Q];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            ok[k] = node_of<D>(kp, (wave * CPT + k) * 64u + lane, x[k], y[k], z[k], c[k]);
            g[k] = 0;
            if (ok[k]) {
                g[k] = geo[c[k]];
                gather<D, S, C, MODE>(src, kp, c[k], f[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < CPT; ++k)
            if (ok[k])
                finish_node<D, S, C, MODE>(f[k], g[k], src, dst, fout, pout, uxo, uyo, uzo, kp, x[k], y[k], z[k],
                                           c[k], pt);
    }
}

// ---------------------------------------------------------------------------
// Batched ensemble: many independent lattices advanced by one launch.  The
// interior nodes of every case are cut into tiles of BLOCK*CPT nodes; a
// block takes a tile (by its index, or from a queue when PERSIST), reads the
// tile's case and runs the fused pull on CPT nodes per thread, strided by
// BLOCK.  Every case keeps its own parameters, populations and node types,
// and the step writes its boundary copies from the case's own push table
// (as the single-lattice step does) unless the case has none.
template <int D, class S, class C, int BLOCK, int CPT, int PERSIST>
__global__ __launch_bounds__(BLOCK) void k_batch_step(const S *__restrict__ src, S *__restrict__ dst,
                                                      const uint8_t *__restrict__ geo,
                                                      const BatchCase *__restrict__ cases,
                                                      const uint32_t *__restrict__ tile_case,
                                                      const uint32_t *__restrict__ tile_first, uint32_t ntiles,
                                                      uint32_t *__restrict__ queue)
{
    __shared__ uint32_t next;
    uint32_t tile = PERSIST ? 0u : blockIdx.x;
    for (;;) {
        if (PERSIST) {
            __syncthreads();  // every thread has read the previous tile index
            if (threadIdx.x == 0) next = atomicAdd(queue, 1u);
            __syncthreads();
            tile = next;
        }
        if (tile >= ntiles) return;  // every block of the queue ends here
        const BatchCase &bc = cases[tile_case[tile]];
        const uint32_t first = tile_first[tile];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            const uint32_t i = first + uint32_t(k * BLOCK) + threadIdx.x;
            int x, y, z;
            uint32_t c;
            if (node_of<D>(bc.kp, i, x, y, z, c))  // cells of a case without pushed copies have no G_PUSH bit
                step_node<D, S, C, K_COLLIDE | K_PUSH>(src + bc.pop_base, dst + bc.pop_base, nullptr, nullptr,
                                                       nullptr, nullptr, nullptr, geo + bc.geo_base, bc.kp, x, y, z,
                                                       c, bc.pt);
        }
        if (!PERSIST) return;
    }
}

// Boundary fill of all cases (slots already offset per case); thread 0 also
// rewinds the tile queue of the step kernel that follows on the stream.
template <class S, class I>
__global__ __launch_bounds__(256) void k_batch_fill(S *__restrict__ buf, const S *__restrict__ prev,
                                                    const I *__restrict__ dslot, const I *__restrict__ sslot,
                                                    int64_t n, uint32_t *__restrict__ queue)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *queue = 0u;
    if (i >= n) return;
    const int64_t s = sslot[i];
    buf[dslot[i]] = s >= 0 ? buf[s] : s == -1 ? Store<S>::template st<float>(0.0f) : prev[-s - 2];
}

// Collision from stored f, p, u (lbm.f90:56-74): the first step after
// InitPDF, and every step in per-phase mode.
template <int D, class S, class C>
__global__ __launch_bounds__(256) void k_collide(const S *__restrict__ fin, S *__restrict__ dst,
                                                 const double *__restrict__ p, const double *__restrict__ ux,
                                                 const double *__restrict__ uy, const double *__restrict__ uz,
                                                 const uint8_t *__restrict__ geo, KParams kp)
{
    using L = Lat<D>;
    int x, y, z;
    uint32_t c;
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x, kp.xcd_remap) * blockDim.x + threadIdx.x;
    if (!node_of<D>(kp, i, x, y, z, c)) return;
    if (!(geo[c] & G_FLUID)) return;
    C f[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; ++q) f[q] = Store<S>::ld(fin[cslot<L::Q>(kp, q, c)]);
    collide<D, C>(f, C(p[c]), C(ux[c]), C(uy[c]), D == 3 ? C(uz[c]) : C(0), kp,
                  [&, srw = sr_word<S>(C(p[c]), c)](int q, C v) {
                      dst[cslot<L::Q>(kp, q, c)] = Store<S>::template st<C>(v, srw + uint32_t(q) * kGolden);
                  });
}

// InitPDF (lbm.f90:29-47): f = feq(p, u) on fluid nodes.
template <int D, class S, class C>
__global__ __launch_bounds__(256) void k_init_pdf(S *__restrict__ fout, const double *__restrict__ p,
                                                  const double *__restrict__ ux, const double *__restrict__ uy,
                                                  const double *__restrict__ uz, const uint8_t *__restrict__ geo,
                                                  KParams kp)
{
    using L = Lat<D>;
    int x, y, z;
    uint32_t c;
    const uint32_t i = xcd_block(blockIdx.x, gridDim.x, kp.xcd_remap) * blockDim.x + threadIdx.x;
    if (!node_of<D>(kp, i, x, y, z, c)) return;
    if (!(geo[c] & G_FLUID)) return;
    const C pp = C(p[c]), u0 = C(ux[c]), u1 = C(uy[c]), u2 = D == 3 ? C(uz[c]) : C(0);
    C udu = u0 * u0 + u1 * u1;
    if (D == 3) udu = udu + u2 * u2;
    const C rho0 = C(kp.rho0);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        C edu = C(0);
        bool first = true;
        auto term = [&](int e, C v) {
            if (e == 0) return;
            if (first) {
                edu = e > 0 ? v : -v;
                first = false;
            } else {
                edu = e > 0 ? edu + v : edu - v;
            }
        };
        term(L::ex(q), u0);
        term(L::ey(q), u1);
        if (D == 3) term(L::ez(q), u2);
        const C feq = C(L::w(q)) * (pp * C(3) + rho0 * ((C(3) * edu + C(4.5) * (edu * edu)) - C(1.5) * udu));
        fout[cslot<L::Q>(kp, q, c)] = Store<S>::template st<C>(feq, sr_word<S>(pp, c) + uint32_t(q) * kGolden);
    }
}

// InitUP (initialization.f90:267-281; initialization_3d.f90:291-307).  The
// perturbation index is the reference's own flat cell id modulo 100.
template <int D>
__global__ __launch_bounds__(256) void k_init_up(double *__restrict__ p, double *__restrict__ ux,
                                                 double *__restrict__ uy, double *__restrict__ uz,
                                                 const uint8_t *__restrict__ geo, const double *__restrict__ rv,
                                                 KParams kp)
{
    int x, y, z;
    uint32_t c;
    if (!decode<D>(kp, blockIdx.x * blockDim.x + threadIdx.x, x, y, z, c)) return;
    const int64_t rid = int64_t(x + 3) + kp.rdy * (y + 3) + (D == 3 ? kp.rdz * (z + 3) : 0);
    const double r = 1.0 + (rv[rid % 100] - 0.5) * 2.0 * 0.2;
    const double g = double(geo[c] & G_FLUID);
    p[c] = 0.0;
    if (D == 2) {
        const double yy = double(y);
        ux[c] = kp.uu * 4.0 * yy * (double(kp.ny) - yy) / double(kp.ny * kp.ny) * g * r;
        uy[c] = 0.0;
    } else {
        const double zz = double(z);
        ux[c] = kp.uu * 4.0 * zz * (double(kp.nz) - zz) / double(kp.nz * kp.nz) * r * g;
        uy[c] = 0.0;
        uz[c] = 0.0;
    }
}

// Fields of a 2D lattice from an FP32 stack f[3][H][W] (p, u_x, u_y at node
// rows y < H and columns x < W; nodes beyond take the last row and column):
// p = ps f0, u_x = us f1 r, u_y = us f2, with r InitUP's perturbation
// factor when rv is given (1 otherwise), and 0 on body nodes.
__global__ __launch_bounds__(256) void k_init_given(double *__restrict__ p, double *__restrict__ ux,
                                                    double *__restrict__ uy, const uint8_t *__restrict__ geo,
                                                    const double *__restrict__ rv, KParams kp,
                                                    const float *__restrict__ f, int H, int W, double ps, double us)
{
    int x, y, z;
    uint32_t c;
    if (!decode<2>(kp, blockIdx.x * blockDim.x + threadIdx.x, x, y, z, c)) return;
    const int64_t rid = int64_t(x + 3) + kp.rdy * (y + 3);
    const double r = rv ? 1.0 + (rv[rid % 100] - 0.5) * 2.0 * 0.2 : 1.0;
    const double g = double(geo[c] & G_FLUID);
    const int64_t k = int64_t(y < H ? y : H - 1) * W + (x < W ? x : W - 1);
    const int64_t hw = int64_t(H) * W;
    p[c] = ps * double(f[k]) * g;
    ux[c] = us * double(f[hw + k]) * r * g;
    uy[c] = us * double(f[2 * hw + k]) * g;
}

// out[y W + x] = 1 on body nodes, 0 on fluid nodes, for nodes y < H, x < W
// of a 2D lattice
__global__ __launch_bounds__(256) void k_body_mask(float *__restrict__ out, const uint8_t *__restrict__ geo,
                                                   KParams kp, int H, int W)
{
    int x, y, z;
    uint32_t c;
    if (!decode<2>(kp, blockIdx.x * blockDim.x + threadIdx.x, x, y, z, c)) return;
    if (x < W && y < H) out[int64_t(y) * W + x] = (geo[c] & G_FLUID) ? 0.0f : 1.0f;
}

// Boundary fill of this step's buffer: dst slot <- this step's fluid slot
// (s >= 0), 0 (s == -1), or a body-cell slot of the previous step's buffer
// (s <= -2, slot -s-2).  Sources are never destinations, so the order of
// the entries does not matter; they are sorted by destination, so a wave's
// writes land on consecutive slots.  I: int32_t or int64_t slot tables.
template <class S, class I>
__global__ __launch_bounds__(256) void k_fill(S *__restrict__ buf, const S *__restrict__ prev,
                                              const I *__restrict__ dslot, const I *__restrict__ sslot, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t s = sslot[i];
    buf[dslot[i]] = s >= 0 ? buf[s] : s == -1 ? Store<S>::template st<float>(0.0f) : prev[-s - 2];
}

// Monitor (lbm.f90:212-217): max of ux^2+uy^2 over fluid nodes except the
// first fluid node in scan order; sqrt is taken on the host (monotone, so
// the maximum of the roots equals the root of the maximum).  A bounded grid
// strides over the nodes with MON_UNROLL independent loads in flight per
// thread; one 64-bit atomicMax per block (non-negative doubles order like
// their bit patterns) into out[0].  Nodes where ux^2+uy^2 is not finite
// (NaN or infinite: a diverged lattice) are left out of the maximum and
// counted into out[1], one atomicAdd per block that found any.
constexpr unsigned MON_BLOCKS = 1024;
constexpr int MON_UNROLL = 4;

template <int D>
__global__ __launch_bounds__(256) void k_monitor(const double *__restrict__ ux, const double *__restrict__ uy,
                                                 const uint8_t *__restrict__ geo, uint32_t skip, KParams kp,
                                                 unsigned long long *__restrict__ out)
{
    __shared__ double red[4];
    __shared__ unsigned bad_blk;
    if (threadIdx.x == 0) bad_blk = 0u;
    __syncthreads();
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double v = 0.0;
    unsigned bad = 0u;
    for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; base < kp.N;
         base += MON_UNROLL * stride) {
        double a[MON_UNROLL];
#pragma unroll
        for (int k = 0; k < MON_UNROLL; ++k) {
            const uint64_t i = base + uint64_t(k) * stride;
            int x, y, z;
            uint32_t c;
            a[k] = 0.0;
            if (i < kp.N && decode<D>(kp, uint32_t(i), x, y, z, c) && (geo[c] & G_FLUID) && c != skip)
                a[k] = ux[c] * ux[c] + uy[c] * uy[c];
        }
#pragma unroll
        for (int k = 0; k < MON_UNROLL; ++k) {
            const bool fin = __builtin_isfinite(a[k]);
            bad += fin ? 0u : 1u;
            v = fin ? fmax(v, a[k]) : v;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        v = fmax(v, __shfl_xor(v, o, 64));
        bad += __shfl_xor(bad, o, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[wave] = v;
        if (bad) atomicAdd(&bad_blk, bad);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < int(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
        atomicMax(out, (unsigned long long)__double_as_longlong(m));
        if (bad_blk) atomicAdd(out + 1, (unsigned long long)bad_blk);
    }
}

// clears the Monitor's two result words (maximum, non-finite count)
__global__ void k_zero_u64(unsigned long long *p)
{
    if (threadIdx.x < 2) p[threadIdx.x] = 0ull;
}

// Mass check: sum over fluid nodes of sum_q f_q (the density deviation), in
// FP64 from every storage form.  Each thread adds its nodes (q in order,
// grid-stride over the nodes), each block writes one partial, and
// k_sum_partials adds the partials in block order: the result depends on
// the data only.
template <int D, class S>
__global__ __launch_bounds__(256) void k_mass(const S *__restrict__ f, const uint8_t *__restrict__ geo, KParams kp,
                                              double *__restrict__ partial)
{
    using L = Lat<D>;
    __shared__ double red[4];
    const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
    double v = 0.0;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < kp.N; i += stride) {
        int x, y, z;
        uint32_t c;
        if (!decode<D>(kp, uint32_t(i), x, y, z, c) || !(geo[c] & G_FLUID)) continue;
        double m = 0.0;
#pragma unroll
        for (int q = 0; q < L::Q; ++q) m += double(Store<S>::ld(f[cslot<L::Q>(kp, q, c)]));
        v += m;
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = red[0];
        for (int w = 1; w < int(blockDim.x >> 6); ++w) t += red[w];
        partial[blockIdx.x] = t;
    }
}

__global__ void k_sum_partials(const double *__restrict__ partial, int64_t n, double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double t = 0.0;
    for (int64_t b = 0; b < n; ++b) t += partial[b];
    out[0] = t;
}

// Momentum exchange on the body links (fluid x -> solid x+e_q) of the
// post-BoundaryCondition buffer: F += e_q (fb(x,q) + fb(x+e_q, opposite q)),
// FP64.  Stage 1 writes one partial sum per block; stage 2 adds the partials
// in block order, so the result does not depend on scheduling.
template <int D, class S>
__global__ __launch_bounds__(256) void k_forces(const S *__restrict__ buf, const int64_t *__restrict__ out_slot,
                                                const int64_t *__restrict__ back_slot,
                                                const int8_t *__restrict__ lq, int64_t n,
                                                double *__restrict__ partial)
{
    using L = Lat<D>;
    __shared__ double red[4][3];
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < n) {
        const int q = lq[i];
        const double m = double(Store<S>::ld(buf[out_slot[i]])) + double(Store<S>::ld(buf[back_slot[i]]));
        s[0] = double(L::ex(q)) * m;
        s[1] = double(L::ey(q)) * m;
        s[2] = double(L::ez(q)) * m;
    }
    for (int d = 0; d < 3; ++d)
        for (int o = 32; o > 0; o >>= 1) s[d] += __shfl_xor(s[d], o, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int d = 0; d < 3; ++d) red[wave][d] = s[d];
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < int(blockDim.x >> 6); ++w) t += red[w][threadIdx.x];
        partial[int64_t(blockIdx.x) * 3 + threadIdx.x] = t;
    }
}

__global__ void k_forces_final(const double *__restrict__ partial, int64_t nblk, double *__restrict__ out)
{
    if (threadIdx.x >= 3) return;
    double t = 0.0;
    for (int64_t b = 0; b < nblk; ++b) t += partial[b * 3 + threadIdx.x];
    out[threadIdx.x] = t;
}

// ---------------------------------------------------------------------------
// host-side dispatch

static inline unsigned nblocks(uint64_t n) { return unsigned((n + 255) / 256); }
// threads of the kernels that use node_of()
static inline uint64_t node_threads(const KParams &kp) { return IMEXLB_CELLMAP ? uint64_t(kp.ncell) : uint64_t(kp.N); }

template <class F>
static hipError_t dispatch_prec(int prec, F &&f)
{
    switch (prec) {
    case 0: return f((double *)nullptr, (double *)nullptr);
    case 1: return f((float *)nullptr, (float *)nullptr);
    case 2: return f((_Float16 *)nullptr, (float *)nullptr);
    case 3: return f((bf16_t *)nullptr, (float *)nullptr);
    case 4: return f((bf16sr_t *)nullptr, (float *)nullptr);
    default: return hipErrorInvalidValue;
    }
}

template <int D, int MODE>
static hipError_t step_dm(int prec, const StepArgs &a, hipStream_t s)
{
    return dispatch_prec(prec, [&](auto *sp, auto *cp) {
        using S = std::remove_pointer_t<decltype(sp)>;
        using C = std::remove_pointer_t<decltype(cp)>;
        constexpr unsigned block = step_block<D, S, MODE>();  // the kernel's launch bound
        constexpr uint64_t per_block = uint64_t(block) * step_cpt<MODE>();
        const unsigned nb = unsigned((node_threads(a.kp) + per_block - 1) / per_block);
        hipLaunchKernelGGL((k_step<D, S, C, MODE>), dim3(nb), dim3(block), 0, s, (const S *)a.src,
                           (S *)a.dst, (S *)a.fout, a.p, a.ux, a.uy, a.uz, a.geo, a.kp, a.push);
        return hipGetLastError();
    });
}

template <int D>
static hipError_t step_d(int prec, int mode, const StepArgs &a, hipStream_t s)
{
    switch (mode) {
    case K_COLLIDE: return step_dm<D, K_COLLIDE>(prec, a, s);
    case K_COLLIDE | K_PUSH: return step_dm<D, K_COLLIDE | K_PUSH>(prec, a, s);
    case K_COLLIDE | K_PUSH | K_IDX64: return step_dm<D, K_COLLIDE | K_PUSH | K_IDX64>(prec, a, s);
    case K_COLLIDE | K_PUSH | K_SMALL:  // the small-lattice form is 2D only
        if constexpr (D == 2) return step_dm<D, K_COLLIDE | K_PUSH | K_SMALL>(prec, a, s);
        return hipErrorInvalidValue;
    case K_WRITE_F: return step_dm<D, K_WRITE_F>(prec, a, s);
    case K_WRITE_F | K_WRITE_PU: return step_dm<D, K_WRITE_F | K_WRITE_PU>(prec, a, s);
    case K_NOSHIFT | K_WRITE_PU: return step_dm<D, K_NOSHIFT | K_WRITE_PU>(prec, a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_step(int dim, int prec, int mode, const StepArgs &a, hipStream_t s)
{
    return dim == 2 ? step_d<2>(prec, mode, a, s) : step_d<3>(prec, mode, a, s);
}

hipError_t launch_collide(int dim, int prec, const StepArgs &a, hipStream_t s)
{
    auto go = [&](auto dtag) {
        constexpr int D = decltype(dtag)::value;
        return dispatch_prec(prec, [&](auto *sp, auto *cp) {
            using S = std::remove_pointer_t<decltype(sp)>;
            using C = std::remove_pointer_t<decltype(cp)>;
            hipLaunchKernelGGL((k_collide<D, S, C>), dim3(nblocks(node_threads(a.kp))), dim3(256), 0, s, (const S *)a.src,
                               (S *)a.dst, a.p, a.ux, a.uy, a.uz, a.geo, a.kp);
            return hipGetLastError();
        });
    };
    return dim == 2 ? go(std::integral_constant<int, 2>{}) : go(std::integral_constant<int, 3>{});
}

hipError_t launch_init_pdf(int dim, int prec, const StepArgs &a, hipStream_t s)
{
    auto go = [&](auto dtag) {
        constexpr int D = decltype(dtag)::value;
        return dispatch_prec(prec, [&](auto *sp, auto *cp) {
            using S = std::remove_pointer_t<decltype(sp)>;
            using C = std::remove_pointer_t<decltype(cp)>;
            hipLaunchKernelGGL((k_init_pdf<D, S, C>), dim3(nblocks(node_threads(a.kp))), dim3(256), 0, s, (S *)a.fout, a.p,
                               a.ux, a.uy, a.uz, a.geo, a.kp);
            return hipGetLastError();
        });
    };
    return dim == 2 ? go(std::integral_constant<int, 2>{}) : go(std::integral_constant<int, 3>{});
}

hipError_t launch_init_up(int dim, const StepArgs &a, const double *rv, hipStream_t s)
{
    if (dim == 2)
        hipLaunchKernelGGL((k_init_up<2>), dim3(nblocks(a.kp.N)), dim3(256), 0, s, a.p, a.ux, a.uy, a.uz, a.geo, rv,
                           a.kp);
    else
        hipLaunchKernelGGL((k_init_up<3>), dim3(nblocks(a.kp.N)), dim3(256), 0, s, a.p, a.ux, a.uy, a.uz, a.geo, rv,
                           a.kp);
    return hipGetLastError();
}

hipError_t launch_init_given(const StepArgs &a, const double *rv, const float *f, int H, int W, double ps, double us,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_init_given, dim3(nblocks(a.kp.N)), dim3(256), 0, s, a.p, a.ux, a.uy, a.geo, rv, a.kp, f, H,
                       W, ps, us);
    return hipGetLastError();
}

hipError_t launch_body_mask(const StepArgs &a, float *out, int H, int W, hipStream_t s)
{
    hipLaunchKernelGGL(k_body_mask, dim3(nblocks(a.kp.N)), dim3(256), 0, s, out, a.geo, a.kp, H, W);
    return hipGetLastError();
}

hipError_t launch_mass(int dim, int prec, const StepArgs &a, double *partial, double *out, hipStream_t s)
{
    const unsigned nb = std::min(nblocks(a.kp.N), MASS_BLOCKS);
    hipError_t e = dispatch_prec(prec, [&](auto *sp, auto *) {
        using S = std::remove_pointer_t<decltype(sp)>;
        if (dim == 2)
            hipLaunchKernelGGL((k_mass<2, S>), dim3(nb), dim3(256), 0, s, (const S *)a.src, a.geo, a.kp, partial);
        else
            hipLaunchKernelGGL((k_mass<3, S>), dim3(nb), dim3(256), 0, s, (const S *)a.src, a.geo, a.kp, partial);
        return hipGetLastError();
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, s, partial, int64_t(nb), out);
    return hipGetLastError();
}

hipError_t launch_fill(int prec, void *buf, const void *prev, const void *dslot, const void *sslot, bool idx32,
                       int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    return dispatch_prec(prec, [&](auto *sp, auto *) {
        using S = std::remove_pointer_t<decltype(sp)>;
        if (idx32)
            hipLaunchKernelGGL((k_fill<S, int32_t>), dim3(nblocks(uint64_t(n))), dim3(256), 0, s, (S *)buf,
                               (const S *)prev, (const int32_t *)dslot, (const int32_t *)sslot, n);
        else
            hipLaunchKernelGGL((k_fill<S, int64_t>), dim3(nblocks(uint64_t(n))), dim3(256), 0, s, (S *)buf,
                               (const S *)prev, (const int64_t *)dslot, (const int64_t *)sslot, n);
        return hipGetLastError();
    });
}

hipError_t launch_forces(int dim, int prec, const void *buf, const int64_t *out_slot, const int64_t *back_slot,
                         const int8_t *lq, int64_t n, double *partial, double *out, hipStream_t s)
{
    const unsigned nb = n > 0 ? nblocks(uint64_t(n)) : 1u;
    hipError_t e = dispatch_prec(prec, [&](auto *sp, auto *) {
        using S = std::remove_pointer_t<decltype(sp)>;
        if (dim == 2)
            hipLaunchKernelGGL((k_forces<2, S>), dim3(nb), dim3(256), 0, s, (const S *)buf, out_slot, back_slot, lq, n,
                               partial);
        else
            hipLaunchKernelGGL((k_forces<3, S>), dim3(nb), dim3(256), 0, s, (const S *)buf, out_slot, back_slot, lq, n,
                               partial);
        return hipGetLastError();
    });
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_forces_final, dim3(1), dim3(64), 0, s, partial, int64_t(nb), out);
    return hipGetLastError();
}

hipError_t launch_monitor(int dim, const StepArgs &a, uint32_t skip, unsigned long long *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, out);
    const unsigned nb = std::min(nblocks(a.kp.N), MON_BLOCKS);
    if (dim == 2)
        hipLaunchKernelGGL((k_monitor<2>), dim3(nb), dim3(256), 0, s, a.ux, a.uy, a.geo, skip, a.kp, out);
    else
        hipLaunchKernelGGL((k_monitor<3>), dim3(nb), dim3(256), 0, s, a.ux, a.uy, a.geo, skip, a.kp, out);
    return hipGetLastError();
}

}  // namespace imexlb

namespace imexlb {

template <class S, class C, int BLOCK, int CPT, int PERSIST>
static void batch_go(unsigned grid, const void *src, void *dst, const uint8_t *geo, const BatchCase *cases,
                     const uint32_t *tc, const uint32_t *tf, uint32_t ntiles, uint32_t *queue, hipStream_t s)
{
    hipLaunchKernelGGL((k_batch_step<2, S, C, BLOCK, CPT, PERSIST>), dim3(grid), dim3(BLOCK), 0, s, (const S *)src,
                       (S *)dst, geo, cases, tc, tf, ntiles, queue);
}

template <class S, class C, int BLOCK, int CPT>
static hipError_t batch_p(int persist, unsigned grid, const void *src, void *dst, const uint8_t *geo,
                          const BatchCase *cases, const uint32_t *tc, const uint32_t *tf, uint32_t ntiles,
                          uint32_t *queue, hipStream_t s)
{
    if (persist)
        batch_go<S, C, BLOCK, CPT, 1>(grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    else
        batch_go<S, C, BLOCK, CPT, 0>(ntiles, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    return hipGetLastError();
}

template <class S, class C, int BLOCK>
static hipError_t batch_c(int cpt, int persist, unsigned grid, const void *src, void *dst, const uint8_t *geo,
                          const BatchCase *cases, const uint32_t *tc, const uint32_t *tf, uint32_t ntiles,
                          uint32_t *queue, hipStream_t s)
{
    switch (cpt) {
    case 1: return batch_p<S, C, BLOCK, 1>(persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    case 2: return batch_p<S, C, BLOCK, 2>(persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    case 4: return batch_p<S, C, BLOCK, 4>(persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    default: return hipErrorInvalidValue;
    }
}

template <class S, class C>
static hipError_t batch_b(int block, int cpt, int persist, unsigned grid, const void *src, void *dst,
                          const uint8_t *geo, const BatchCase *cases, const uint32_t *tc, const uint32_t *tf,
                          uint32_t ntiles, uint32_t *queue, hipStream_t s)
{
    switch (block) {
    case 64: return batch_c<S, C, 64>(cpt, persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    case 128: return batch_c<S, C, 128>(cpt, persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    case 256: return batch_c<S, C, 256>(cpt, persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    case 512: return batch_c<S, C, 512>(cpt, persist, grid, src, dst, geo, cases, tc, tf, ntiles, queue, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_step(int prec, int block, int cpt, int persist, unsigned grid, const void *src, void *dst,
                             const uint8_t *geo, const BatchCase *cases, const uint32_t *tile_case,
                             const uint32_t *tile_first, uint32_t ntiles, uint32_t *queue, hipStream_t s)
{
    if (ntiles == 0) return hipSuccess;
    switch (prec) {
    case 0:
        return batch_b<double, double>(block, cpt, persist, grid, src, dst, geo, cases, tile_case, tile_first, ntiles,
                                       queue, s);
    case 1:
        return batch_b<float, float>(block, cpt, persist, grid, src, dst, geo, cases, tile_case, tile_first, ntiles,
                                     queue, s);
    case 2:
        return batch_b<_Float16, float>(block, cpt, persist, grid, src, dst, geo, cases, tile_case, tile_first, ntiles,
                                        queue, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batch_fill(int prec, void *buf, const void *prev, const void *dslot, const void *sslot,
                             bool idx32, int64_t n, uint32_t *queue, hipStream_t s)
{
    const unsigned nb = n > 0 ? nblocks(uint64_t(n)) : 1u;
    return dispatch_prec(prec, [&](auto *sp, auto *) {
        using S = std::remove_pointer_t<decltype(sp)>;
        if (idx32)
            hipLaunchKernelGGL((k_batch_fill<S, int32_t>), dim3(nb), dim3(256), 0, s, (S *)buf, (const S *)prev,
                               (const int32_t *)dslot, (const int32_t *)sslot, n, queue);
        else
            hipLaunchKernelGGL((k_batch_fill<S, int64_t>), dim3(nb), dim3(256), 0, s, (S *)buf, (const S *)prev,
                               (const int64_t *)dslot, (const int64_t *)sslot, n, queue);
        return hipGetLastError();
    });
}

}  // namespace imexlb
