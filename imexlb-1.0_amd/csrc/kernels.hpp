// kernels.hpp -- launch interface between the context (capi.cpp) and the
// gfx950 kernels (lbm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace imexlb {

// n / d for 32-bit n by multiply-high and shift (round-up method):
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1, q = (mulhi(n,m) + n) >> s.
struct FastDiv {
    uint32_t m = 1, s = 0;
    void init(uint32_t d)
    {
        s = 0;
        while ((uint64_t(1) << s) < d) ++s;
        m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << s) - d)) / d + 1);
    }
    uint32_t host_div(uint32_t n) const
    {
        const uint64_t hi = (uint64_t(n) * m) >> 32;
        return uint32_t((hi + n) >> s);
    }
};

// Population layout: 0 = one plane per direction; T > 0 = tiles of T padded
// cells, all directions of a tile contiguous (slot ((c/T)*Q + q)*T + c%T).
// Tiles of 64 cells are the default: one wave's 27 (or 9) streams then fall
// in one 13.8 KB (FP64 D3Q27) stretch of memory instead of 27 planes, which
// the DRAM pages serve faster (5.3 vs 4.4 TB/s for a copy of that shape).
#ifndef IMEXLB_TILE
#define IMEXLB_TILE 64
#endif
static_assert(IMEXLB_TILE >= 0 && (IMEXLB_TILE & (IMEXLB_TILE - 1)) == 0, "tile size: 0 or a power of two");

// Thread i of the step kernels handles padded cell i (1; threads on ghost
// cells leave, and a wave covers one tile) or interior node i (0).
#ifndef IMEXLB_CELLMAP
#define IMEXLB_CELLMAP 1
#endif

// Geometry and scalars every kernel needs (passed by value).
struct KParams {
    int nx, ny, nz;      // largest interior node index per axis
    uint32_t N;          // interior nodes (nx+1)(ny+1)(nz+1)
    FastDiv div_x;       // / (nx+1)
    FastDiv div_y;       // / (ny+1)
    int64_t o0;          // padded index of interior node (0,0,0)
    int64_t Px, PxPy;    // padded strides of one y / z step
    int64_t plane;       // values per population plane
    int64_t rdy, rdz;    // reference-layout strides (InitUP perturbation index)
    double ome, rho0, uu;
    uint32_t ncell;      // padded cells Px*Py*Pz
    int64_t Py;          // padded rows per z plane
    FastDiv div_px;      // / Px
    FastDiv div_py;      // / Py
    double inv_rho0;     // 1/rho0 (FP32-compute forms multiply by it)
    uint32_t xcd_remap;  // block order over the XCDs: 0 hardware, 1 an eighth each, R >= 2 runs of R
    float ome_f, rho0_f, uu_f, inv_rho0_f;  // the same scalars rounded to FP32 once
    uint32_t back_tiles; // ceil((1 + Px + PxPy) / 64): the farthest tile a pull reads behind its own
};

// Position of direction q inside a tile.  IMEXLB_QGROUP = 1 stores the
// directions that pull from the same source row (equal e_y, e_z) next to
// each other, e_x fastest, so a wave reads each source tile as one run;
// 0 keeps the lattice numbering (initialization.f90:112-123).
#ifndef IMEXLB_QGROUP
#define IMEXLB_QGROUP 0
#endif
__host__ __device__ constexpr int store_q(int q, int Q)
{
    return !IMEXLB_QGROUP ? q
           : Q == 9       ? (q % 3) * 3 + q / 3
                          : ((q / 9) * 3 + q % 3) * 3 + (q % 9) / 3;
}

// slot of direction q at padded cell c, host side (kernels use cslot)
inline int64_t pop_slot(int64_t q, int64_t c, int64_t plane, int Q)
{
#if IMEXLB_TILE
    (void)plane;
    return ((c / IMEXLB_TILE) * Q + store_q(int(q), Q)) * IMEXLB_TILE + c % IMEXLB_TILE;
#else
    (void)Q;
    return q * plane + c;
#endif
}

// K_PUSH: the collide step also writes the boundary copies of its output
// (PushTable), so the buffer needs no boundary fill; K_IDX64: the push
// table's destination slots are int64 (else int32); K_SMALL: a small,
// latency-bound 2D lattice -- blocks of one wave, each a 64-cell tile whose
// boundary copies are spread over its lanes through LDS, the output written
// with ordinary (cached) stores (with the XCD-contiguous block order each
// XCD keeps stepping the same eighth of the lattice, in its own L2).
enum : int { K_COLLIDE = 1, K_WRITE_F = 2, K_WRITE_PU = 4, K_NOSHIFT = 8, K_PUSH = 16, K_IDX64 = 32, K_SMALL = 64 };
constexpr unsigned kSmallBlock = 64;

// node-type byte on the device: fluid node; node with boundary copies to push
enum : uint8_t { G_FLUID = 1, G_PUSH = 2 };

// The resolved boundary table grouped by source node: entries beg[c] ..
// beg[c+1]-1 of padded cell c copy its direction q[e] to slot dst[e] of the
// same buffer -- the post-collision value of a fluid node, or for a body
// cell the previous step's value (from the other buffer).
struct PushTable {
    const int32_t *beg = nullptr;
    const uint8_t *q = nullptr;
    const void *dst = nullptr;  // int32 or int64 slots (K_IDX64)
    // q * 64 + (source cell % 64) of each entry, where the small-lattice
    // step and the 16/32-bit D3Q27 step find the copied value in their
    // wave's LDS
    const uint16_t *lds = nullptr;
};

// One case of a batched ensemble: its parameters and where its populations
// and node types start in the shared arrays.
struct BatchCase {
    KParams kp;
    int64_t pop_base;  // first population slot of the case
    int64_t geo_base;  // first cell of the case (node types, p, u)
    PushTable pt;      // the case's boundary copies (int32 slots within the case), pushed by the step
};

struct StepArgs {
    KParams kp;
    const void *src = nullptr;  // population input
    void *dst = nullptr;        // post-collision output
    void *fout = nullptr;       // array f output
    double *p = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr;
    const uint8_t *geo = nullptr;
    PushTable push;
};

// prec: 0 FP64, 1 FP32, 2 FP16 storage + FP32 arithmetic, 3 BF16 storage + FP32 arithmetic
hipError_t launch_step(int dim, int prec, int mode, const StepArgs &a, hipStream_t s);
hipError_t launch_collide(int dim, int prec, const StepArgs &a, hipStream_t s);
hipError_t launch_init_pdf(int dim, int prec, const StepArgs &a, hipStream_t s);
hipError_t launch_init_up(int dim, const StepArgs &a, const double *rv, hipStream_t s);
// 2D fields from an FP32 device stack [3][H][W] (p, u_x, u_y), scaled by
// ps / us, u_x times InitUP's perturbation when rv is given, 0 on bodies
hipError_t launch_init_given(const StepArgs &a, const double *rv, const float *f, int H, int W, double ps, double us,
                             hipStream_t s);
// sum over fluid nodes of sum_q f_q of the populations a.src, FP64, into
// out[0]; partial holds MASS_BLOCKS doubles
constexpr unsigned MASS_BLOCKS = 1024;
hipError_t launch_mass(int dim, int prec, const StepArgs &a, double *partial, double *out, hipStream_t s);
// out[y W + x] = 1 on body nodes, 0 on fluid nodes (2D, y < H, x < W)
hipError_t launch_body_mask(const StepArgs &a, float *out, int H, int W, hipStream_t s);
// Boundary fill; the slot tables hold int32 (idx32) or int64 entries.
hipError_t launch_fill(int prec, void *buf, const void *prev, const void *dslot, const void *sslot, bool idx32,
                       int64_t n, hipStream_t s);
hipError_t launch_monitor(int dim, const StepArgs &a, uint32_t skip, unsigned long long *out, hipStream_t s);
// Batched ensemble (D2Q9; prec 0, 1 or 2).  block in {64,128,256,512},
// cpt in {1,2,4}; persist launches `grid` blocks that take tiles from
// `queue` (rewound by launch_batch_fill), otherwise one block per tile.
hipError_t launch_batch_step(int prec, int block, int cpt, int persist, unsigned grid, const void *src, void *dst,
                             const uint8_t *geo, const BatchCase *cases, const uint32_t *tile_case,
                             const uint32_t *tile_first, uint32_t ntiles, uint32_t *queue, hipStream_t s);
hipError_t launch_batch_fill(int prec, void *buf, const void *prev, const void *dslot, const void *sslot,
                             bool idx32, int64_t n, uint32_t *queue, hipStream_t s);
// body-link momentum exchange; partial holds 3 doubles per 256 links (at least 3)
hipError_t launch_forces(int dim, int prec, const void *buf, const int64_t *out_slot, const int64_t *back_slot,
                         const int8_t *lq, int64_t n, double *partial, double *out, hipStream_t s);

}  // namespace imexlb
