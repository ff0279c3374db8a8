// context.hpp -- internal: the lattice context behind the C ABI, shared by
// capi.cpp (single lattice) and batch.cpp (batched ensemble).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/imexlb.h"
#include "geometry.hpp"
#include "kernels.hpp"

namespace imexlb {

inline std::string &last_error()
{
    thread_local std::string s;
    return s;
}

inline int fail(int code, const std::string &msg)
{
    last_error() = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? IMEXLB_E_OOM : IMEXLB_E_HIP,                       \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                            \
    } while (0)

enum Stage { S_NONE, S_PRE, S_COLLIDED, S_FILLED, S_PENDING, S_STREAMED, S_FUSED };

struct TimedLaunch {
    int cls;
    hipEvent_t a, b;
};


}  // namespace imexlb

using namespace imexlb;

struct imexlb_ctx {
    int dim = 2, Q = 9, device = -1, prec = 0;
    bool host_only = true;
    RefGeometry g;
    BcTable bc;
    bool geom_set = false;
    double ome = 0.0, rho0 = 1.0, uu = 0.0;

    // device layout
    int64_t Px = 0, Py = 0, Pz = 1, PxPy = 0, o0 = 0, plane = 0;
    size_t ssz = 8;
    KParams kp{};
    uint32_t skip_c = 0xffffffffu;

    hipStream_t stream = nullptr;
    void *fb[2] = {nullptr, nullptr};
    void *f = nullptr;
    double *p = nullptr, *ux = nullptr, *uy = nullptr, *uz = nullptr, *rv = nullptr;
    uint8_t *geo = nullptr;
    void *fill_dst = nullptr, *fill_src = nullptr;  // int32 (fill32) or int64 slots
    bool fill32 = false;
    // boundary copies pushed by the fused step (grouped resolved table)
    int32_t *push_beg = nullptr;
    uint8_t *push_q = nullptr;
    void *push_dst = nullptr;   // int32 (fill32) or int64 slots
    uint16_t *push_lds = nullptr;  // LDS position of each entry's value
    bool push_enabled = true;   // IMEXLB_PUSH=0 keeps the separate fill
    bool push_on = false;       // enabled and the table qualifies
    bool prefilled = false;     // buffer cur already holds its boundary copies
    int64_t nfill = 0;
    unsigned long long *red = nullptr;
    // body links for the momentum-exchange force
    int64_t *link_out = nullptr, *link_back = nullptr;
    int8_t *link_q = nullptr;
    int64_t nlink = 0;
    double *force_partial = nullptr, *force_out = nullptr;
    double *mass_partial = nullptr;  // MASS_BLOCKS partial sums, then the total
    size_t dev_bytes = 0;
    // test hook: IMEXLB_FAIL_ALLOC=N (read at creation) makes this
    // context's N-th device allocation fail with IMEXLB_E_OOM
    int64_t alloc_calls = 0, fail_alloc_at = 0;

    // Forget the geometry after a failed SetGeometry: every call that needs
    // one (InitUP, the phases, Monitor, downloads) then returns
    // IMEXLB_E_STATE until SetGeometry succeeds.
    void forget_geometry()
    {
        geom_set = false;
        nfill = nlink = 0;
        push_on = prefilled = false;
        stage = S_NONE;
        have_pu = have_f = have_fb = false;
        lazy_src = -1;
        drop_graphs();
    }

    // phase state
    Stage stage = S_NONE;
    int cur = 0;
    bool fused = true;
    bool have_pu = false, have_f = false, have_fb = false;
    int lazy_src = -1;

    // replay of pairs of fused steps as HIP graphs of kGraphPairs[j] pairs,
    // one per size and starting buffer (cur); advance() takes the largest
    // that fit first (longer graphs: fewer gaps between replays, 3 % on the
    // 400x100 cylinder from 16 to 64 pairs)
    bool graphs = true;
    static constexpr int kGraphSizes = 3;
    static constexpr int kGraphPairs[kGraphSizes] = {64, 8, 1};
    hipGraphExec_t graph[kGraphSizes][2] = {};
    void drop_graphs()
    {
        for (auto &set : graph)
            for (int k = 0; k < 2; ++k)
                if (set[k]) {
                    (void)hipGraphExecDestroy(set[k]);
                    set[k] = nullptr;
                }
    }

    // Fused steps asked for through the phase calls (PostProcessing of a
    // lattice whose step writes its own boundary copies) and not launched
    // yet: they run as graph replays once kDeferChunk are pending (one graph
    // of kDeferChunk / 2 pairs, so the GPU works while the caller goes on)
    // and when the state is next needed (any other call that launches work
    // or reads device memory).  Step k of them reads fb[defer_src ^ (k & 1)].
    int64_t deferred = 0;
    int defer_src = 0;
    static constexpr int64_t kDeferChunk = 16;
    bool can_defer() const { return fused && graphs && !prof && push_on; }

    // profiling
    bool prof = false;
    std::vector<TimedLaunch> timed;
    static constexpr size_t kMaxTimed = 512;
    int64_t prof_n[3] = {0, 0, 0};
    double prof_ms[3] = {0, 0, 0};

    int64_t pidx(int x, int y, int z) const { return o0 + x + Px * y + PxPy * z; }

    StepArgs args() const
    {
        StepArgs a;
        a.kp = kp;
        a.p = p;
        a.ux = ux;
        a.uy = uy;
        a.uz = uz;
        a.geo = geo;
        a.push.beg = push_beg;
        a.push.q = push_q;
        a.push.dst = push_dst;
        a.push.lds = push_lds;
        return a;
    }

    int begin(int cls, hipEvent_t &b)
    {
        (void)cls;
        if (!prof) return IMEXLB_OK;
        HIP_TRY(hipEventCreate(&b));
        HIP_TRY(hipEventRecord(b, stream));
        return IMEXLB_OK;
    }
    int end(int cls, hipEvent_t b)
    {
        if (!prof) return IMEXLB_OK;
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        HIP_TRY(hipEventRecord(e, stream));
        timed.push_back(TimedLaunch{cls, b, e});
        // bound the live event pairs of a long timed run: fold the oldest
        // half into the totals (waiting for them) once kMaxTimed are held
        if (timed.size() >= kMaxTimed) return harvest(kMaxTimed / 2);
        return IMEXLB_OK;
    }

    // add the elapsed time of all but the newest `keep` event pairs to
    // prof_n / prof_ms and destroy them (waits for those launches)
    int harvest(size_t keep)
    {
        const size_t n = timed.size() > keep ? timed.size() - keep : 0;
        for (size_t k = 0; k < n; ++k) {
            TimedLaunch &t = timed[k];
            HIP_TRY(hipEventSynchronize(t.b));
            float v = 0.0f;
            HIP_TRY(hipEventElapsedTime(&v, t.a, t.b));
            prof_n[t.cls] += 1;
            prof_ms[t.cls] += double(v);
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        timed.erase(timed.begin(), timed.begin() + std::ptrdiff_t(n));
        return IMEXLB_OK;
    }

    template <class F>
    int launch(int cls, F &&fn)
    {
        hipEvent_t b = nullptr;
        int rc = begin(cls, b);
        if (rc) return rc;
        hipError_t e = fn();
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
        return end(cls, b);
    }

    // BoundaryCondition on buffer `buf`; the other buffer holds the previous
    // step's state (every Collision writes the buffer the previous step did
    // not use)
    int run_fill(int buf)
    {
        return launch(1, [&] {
            return launch_fill(prec, fb[buf], fb[1 - buf], fill_dst, fill_src, fill32, nfill, stream);
        });
    }

    // mode bits of the fused collide step for this lattice (K_SMALL for 2D
    // lattices of at most small_cells padded cells)
    int collide_mode() const
    {
        if (!push_on) return K_COLLIDE;
        if (!fill32) return K_COLLIDE | K_PUSH | K_IDX64;
        return K_COLLIDE | K_PUSH | (dim == 2 && int64_t(kp.ncell) <= small_cells ? K_SMALL : 0);
    }
    int64_t small_cells = 0;  // set at creation (IMEXLB_SMALL_CELLS)

    int run_step(int mode, const void *src, void *dst, void *fout)
    {
        StepArgs a = args();
        a.src = src;
        a.dst = dst;
        a.fout = fout;
        return launch((mode & K_COLLIDE) ? 0 : 2, [&] { return launch_step(dim, prec, mode, a, stream); });
    }

    // rebuild f, p, u of the current pre-collision state if the fused
    // kernel skipped writing them
    int materialize()
    {
        if (have_pu && have_f) return IMEXLB_OK;
        if (lazy_src < 0) {
            if (have_pu) return IMEXLB_OK;
            return fail(IMEXLB_E_STATE, "fields requested before InitUP/upload");
        }
        int rc = run_step(K_WRITE_F | K_WRITE_PU, fb[lazy_src], nullptr, f);
        if (rc) return rc;
        have_pu = have_f = true;
        return IMEXLB_OK;
    }

    void release()
    {
        if (host_only) return;
        (void)hipSetDevice(device);
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto &t : timed) {
            (void)hipEventDestroy(t.a);
            (void)hipEventDestroy(t.b);
        }
        timed.clear();
        void *ptrs[] = {fb[0],     fb[1],  f,         p,        ux,       uy,        uz,     rv,
                        geo,       fill_dst, fill_src, red,     link_out, link_back, link_q, force_partial,
                        force_out, push_beg, push_q,   push_dst, push_lds, mass_partial};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
        fb[0] = fb[1] = f = nullptr;
        push_beg = nullptr;
        push_q = nullptr;
        push_dst = nullptr;
        push_lds = nullptr;
        p = ux = uy = uz = rv = nullptr;
        geo = nullptr;
        fill_dst = fill_src = nullptr;
        red = nullptr;
        link_out = link_back = nullptr;
        link_q = nullptr;
        force_partial = force_out = nullptr;
        mass_partial = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace imexlb {

inline int dev_alloc(imexlb_ctx *c, void **ptr, size_t bytes)
{
    if (c->fail_alloc_at > 0 && ++c->alloc_calls == c->fail_alloc_at) {
        *ptr = nullptr;
        return fail(IMEXLB_E_OOM, "hipMalloc(" + std::to_string(bytes) + "): failure injected by IMEXLB_FAIL_ALLOC");
    }
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(e == hipErrorOutOfMemory ? IMEXLB_E_OOM : IMEXLB_E_HIP,
                    "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    c->dev_bytes += bytes;
    e = hipMemset(*ptr, 0, bytes);
    if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return IMEXLB_OK;
}

// Device-layout tables of one lattice, built on the host from its geometry:
// node types, the resolved boundary fill (dst, src in FillEntry encoding
// with device slots) and the body links of the force.
struct DeviceTables {
    std::vector<uint8_t> geo;
    uint32_t skip_c = 0xffffffffu;
    std::vector<int64_t> fill_dst, fill_src;
    std::vector<int32_t> push_beg;  // ncell + 1 offsets
    std::vector<uint8_t> push_q;
    std::vector<int64_t> push_dst;
    std::vector<uint16_t> push_lds;
    bool push_ok = true;  // every copy's source is a fluid node or a body cell
    std::vector<int64_t> link_out, link_back;
    std::vector<int8_t> link_q;
};
int build_device_tables(const imexlb_ctx *c, DeviceTables &t);

// The boundary fill's slot tables fit int32 entries when every slot and its
// previous-step encoding (-slot-2) do.
inline bool fill_fits_32(int64_t nslots) { return nslots < (int64_t(1) << 31) - 2; }

// test hook: IMEXLB_SLOTS64=1 (read when the boundary tables are built)
// stores them with int64 slots also where int32 would do
inline bool force_slots64()
{
    const char *v = std::getenv("IMEXLB_SLOTS64");
    return v && v[0] == '1';
}

// copy a slot table to the device as int32 or int64 entries
inline hipError_t upload_slots(void *dev, const std::vector<int64_t> &v, bool idx32)
{
    if (v.empty()) return hipSuccess;
    if (!idx32) return hipMemcpy(dev, v.data(), v.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    std::vector<int32_t> n(v.begin(), v.end());
    return hipMemcpy(dev, n.data(), n.size() * sizeof(int32_t), hipMemcpyHostToDevice);
}

}  // namespace imexlb
