// batch.cpp -- batched ensemble: many independent D2Q9 lattices of their own
// size and parameters advanced together.  Each case keeps the single-lattice
// layout and boundary tables (built by a host-only context); the cases are
// concatenated in one set of device arrays, so one boundary-fill launch and
// one step launch advance all of them; the step writes the cases' boundary
// copies itself (each case's push table), so after the first step only the
// cases whose table does not qualify need the fill.  The step kernel takes the cases'
// nodes in tiles of block*cells_per_thread nodes, by block index or from a
// queue (persistent blocks); imexlb_batch_tune times the configurations and
// keeps the fastest.  Pairs of steps replay as a HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <new>
#include <vector>

#include "context.hpp"

struct imexlb_batch {
    int ncase = 0, prec = 0, device = 0;
    size_t ssz = 8;
    std::vector<imexlb_ctx *> host;  // host-only contexts: geometry, tables, layout
    std::vector<BatchCase> cases;
    std::vector<int64_t> link_base, link_count;
    int64_t pop_total = 0, cell_total = 0, node_total = 0, max_links = 0;

    hipStream_t stream = nullptr;
    void *fb[2] = {nullptr, nullptr};
    void *f = nullptr;
    double *p = nullptr, *ux = nullptr, *uy = nullptr, *rv = nullptr;
    uint8_t *geo = nullptr;
    void *fill_dst = nullptr, *fill_src = nullptr;  // int32 (fill32) or int64 slots
    bool fill32 = false;
    int64_t nfill = 0;
    // the fill's first nfill_step entries belong to cases without pushed
    // copies and run every step; all nfill run before the first step
    int64_t nfill_step = 0;
    bool need_fill = true;
    int32_t *push_beg = nullptr, *push_dst = nullptr;  // every case's push table, concatenated
    uint8_t *push_q = nullptr;
    int64_t *link_out = nullptr, *link_back = nullptr;
    int8_t *link_q = nullptr;
    double *force_partial = nullptr, *force_out = nullptr;
    unsigned long long *mon = nullptr;  // Monitor's two result words per case
    std::vector<uint32_t> skip;         // per case: the fluid node Monitor leaves out
    BatchCase *d_cases = nullptr;
    uint32_t *queue = nullptr;
    uint32_t *tile_case = nullptr, *tile_first = nullptr;
    uint32_t ntiles = 0;
    size_t dev_bytes = 0;

    int block = 256, cpt = 1, persist = 0;
    unsigned persist_grid = 1024;
    // captured pairs of steps, kGraphPairs[j] pairs per graph, by size and
    // starting buffer; advance replays the larger first
    static constexpr int kGraphSizes = 2;
    static constexpr int kGraphPairs[kGraphSizes] = {16, 1};
    hipGraphExec_t graph[kGraphSizes][2] = {};
    bool graphs = true;  // replay captured pairs of steps (IMEXLB_GRAPHS=0: launch every step)

    int cur = 1, lazy = -1;
    bool initialized = false, have_pu = false;

    void drop_graphs()
    {
        for (auto &set : graph)
            for (auto &g : set)
                if (g) {
                    (void)hipGraphExecDestroy(g);
                    g = nullptr;
                }
    }

    void release()
    {
        if (stream) {
            (void)hipSetDevice(device);
            (void)hipStreamSynchronize(stream);
        }
        drop_graphs();
        void *ptrs[] = {fb[0],     fb[1],      f,        p,         ux,        uy,            rv,
                        geo,       fill_dst,   fill_src, link_out,  link_back, link_q,        force_partial,
                        force_out, mon,        d_cases,  queue,     tile_case, tile_first,
                        push_beg,  push_dst,   push_q};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
        for (imexlb_ctx *c : host) imexlb_destroy(c);
        host.clear();
    }

    // one case's kernel arguments, pointers offset to the case's arrays
    StepArgs case_args(int i) const
    {
        StepArgs a;
        a.kp = cases[size_t(i)].kp;
        const int64_t cb = cases[size_t(i)].geo_base;
        a.p = p + cb;
        a.ux = ux + cb;
        a.uy = uy + cb;
        a.uz = nullptr;
        a.geo = geo + cb;
        return a;
    }
    void *pop(void *base, int i) const { return (uint8_t *)base + size_t(cases[size_t(i)].pop_base) * ssz; }

    int build_tiles()
    {
        const uint32_t T = uint32_t(block * cpt);
        std::vector<uint32_t> tc, tf;
        for (int i = 0; i < ncase; ++i)
            for (uint32_t first = 0; first < (IMEXLB_CELLMAP ? cases[size_t(i)].kp.ncell : cases[size_t(i)].kp.N);
                 first += T) {
                tc.push_back(uint32_t(i));
                tf.push_back(first);
            }
        if (tile_case) (void)hipFree(tile_case);
        if (tile_first) (void)hipFree(tile_first);
        tile_case = tile_first = nullptr;
        ntiles = uint32_t(tc.size());
        HIP_TRY(hipMalloc((void **)&tile_case, std::max<size_t>(tc.size(), 1) * 4));
        HIP_TRY(hipMalloc((void **)&tile_first, std::max<size_t>(tf.size(), 1) * 4));
        if (!tc.empty()) {
            HIP_TRY(hipMemcpy(tile_case, tc.data(), tc.size() * 4, hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(tile_first, tf.data(), tf.size() * 4, hipMemcpyHostToDevice));
        }
        return IMEXLB_OK;
    }

    // one step from buffer b: the boundary fill (every case's entries when
    // full, else those of the cases without pushed copies; with persistent
    // blocks the fill launch also rewinds their tile queue), then the step
    hipError_t enqueue_step(int b, bool full)
    {
        const int64_t n = full ? nfill : nfill_step;
        if (n > 0 || persist) {
            hipError_t e = launch_batch_fill(prec, fb[b], fb[1 - b], fill_dst, fill_src, fill32, n, queue, stream);
            if (e != hipSuccess) return e;
        }
        return launch_batch_step(prec, block, cpt, persist, persist_grid, fb[b], fb[1 - b], geo, d_cases, tile_case,
                                 tile_first, ntiles, queue, stream);
    }

    int one_step()
    {
        hipError_t e = enqueue_step(cur, need_fill);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch step: ") + hipGetErrorString(e));
        need_fill = false;
        lazy = cur;
        cur = 1 - cur;
        have_pu = false;
        return IMEXLB_OK;
    }

    int pair_graph(int size, hipGraphExec_t *out)
    {
        hipGraphExec_t &slot = graph[size][cur];
        if (slot) {
            *out = slot;
            return IMEXLB_OK;
        }
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int k = 0; k < kGraphPairs[size] && e == hipSuccess; ++k) {
            e = enqueue_step(cur, false);
            if (e == hipSuccess) e = enqueue_step(1 - cur, false);
        }
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(IMEXLB_E_HIP, std::string("batch graph capture: ") + hipGetErrorString(e ? e : e2));
        }
        hipGraphExec_t x = nullptr;
        e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch graph: ") + hipGetErrorString(e));
        slot = x;
        *out = x;
        return IMEXLB_OK;
    }

    // f, p, u of case i from the buffer the last step streamed from
    int materialize(int i)
    {
        if (have_pu) return IMEXLB_OK;
        if (lazy < 0) return fail(IMEXLB_E_STATE, "batch fields requested before imexlb_batch_init");
        StepArgs a = case_args(i);
        a.src = pop(fb[lazy], i);
        a.fout = pop(f, i);
        hipError_t e = launch_step(2, prec, K_WRITE_F | K_WRITE_PU, a, stream);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch materialize: ") + hipGetErrorString(e));
        return IMEXLB_OK;
    }
};

namespace {

int balloc(imexlb_batch *b, void **ptr, size_t bytes)
{
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(e == hipErrorOutOfMemory ? IMEXLB_E_OOM : IMEXLB_E_HIP, "batch hipMalloc failed");
    }
    b->dev_bytes += bytes;
    e = hipMemset(*ptr, 0, bytes);
    if (e != hipSuccess) return fail(IMEXLB_E_HIP, "batch hipMemset failed");
    return IMEXLB_OK;
}

#define CHECK_BATCH(b)                                                    \
    do {                                                                  \
        if (!(b)) return fail(IMEXLB_E_ARG, "null batch");               \
        HIP_TRY(hipSetDevice((b)->device));                               \
    } while (0)

}  // namespace

extern "C" {

int imexlb_batch_create(imexlb_batch **out, int ncase, const int32_t *nxny, const double *params, int device,
                        int precision)
{
    return imexlb_batch_create_shapes(out, ncase, nxny, params, nullptr, device, precision);
}

int imexlb_batch_create_shapes(imexlb_batch **out, int ncase, const int32_t *nxny, const double *params,
                               const double *shapes, int device, int precision)
{
    if (!out || !nxny || !params || ncase <= 0) return fail(IMEXLB_E_ARG, "batch arguments");
    *out = nullptr;
    if (precision < IMEXLB_F64 || precision > IMEXLB_F16_F32)
        return fail(IMEXLB_E_ARG, "batch precision must be FP64, FP32 or FP16 storage");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(IMEXLB_E_ARG, "device index out of range");
    std::unique_ptr<imexlb_batch, void (*)(imexlb_batch *)> b(new (std::nothrow) imexlb_batch,
                                                              [](imexlb_batch *x) {
                                                                  if (x) {
                                                                      x->release();
                                                                      delete x;
                                                                  }
                                                              });
    if (!b) return fail(IMEXLB_E_OOM, "batch");
    b->ncase = ncase;
    b->prec = precision;
    b->device = device;
    b->ssz = precision == 0 ? 8 : precision == 1 ? 4 : 2;
    std::vector<int64_t> fd, fs, lo, lbk;
    std::vector<int64_t> fd_push, fs_push;  // fill entries of the cases with pushed copies (first step only)
    std::vector<int32_t> pbeg, pdst;        // push tables of those cases, concatenated
    std::vector<uint8_t> pq;
    std::vector<int64_t> beg_off, ent_off;  // per case (-1: the case is filled every step)
    std::vector<int8_t> lq;
    std::vector<uint8_t> geo;
    const char *pv = std::getenv("IMEXLB_PUSH");
    const bool push_enabled = !(pv && pv[0] == '0');
    for (int i = 0; i < ncase; ++i) {
        imexlb_ctx *c = nullptr;
        int rc = imexlb_create(&c, 2, nxny[2 * i], nxny[2 * i + 1], 0, -1, precision);
        if (rc) return rc;
        b->host.push_back(c);
        c->ssz = b->ssz;
        const double *pr = params + 5 * i;  // pos, radius, ome, rho0, uu
        const double aspect = shapes ? shapes[2 * i] : 1.0, angle = shapes ? shapes[2 * i + 1] : 0.0;
        if ((rc = imexlb_set_params(c, pr[2], pr[3], pr[4])) ||
            (rc = imexlb_set_geometry_shape(c, pr[0], pr[1], aspect, angle)))
            return rc;
        DeviceTables t;
        if ((rc = build_device_tables(c, t))) return rc;
        b->skip.push_back(t.skip_c);
        BatchCase bc;
        bc.kp = c->kp;
        bc.pop_base = b->pop_total;
        bc.geo_base = b->cell_total;
        b->cases.push_back(bc);
        const int64_t pb = bc.pop_base;
        const bool push = push_enabled && t.push_ok && fill_fits_32(int64_t(c->Q) * c->plane);
        std::vector<int64_t> &fdc = push ? fd_push : fd, &fsc = push ? fs_push : fs;
        for (size_t k = 0; k < t.fill_dst.size(); ++k) {
            const int64_t s = t.fill_src[k];
            fdc.push_back(t.fill_dst[k] + pb);
            fsc.push_back(s >= 0 ? s + pb : s == -1 ? -1 : encode_prev(decode_prev(s) + pb));
        }
        if (push) {  // slots within the case: the step offsets dst by pop_base
            beg_off.push_back(int64_t(pbeg.size()));
            ent_off.push_back(int64_t(pq.size()));
            pbeg.insert(pbeg.end(), t.push_beg.begin(), t.push_beg.end());
            pq.insert(pq.end(), t.push_q.begin(), t.push_q.end());
            for (int64_t d : t.push_dst) pdst.push_back(int32_t(d));
        } else {
            beg_off.push_back(-1);
            ent_off.push_back(-1);
            for (uint8_t &g : t.geo) g = uint8_t(g & ~G_PUSH);
        }
        b->link_base.push_back(int64_t(lo.size()));
        b->link_count.push_back(int64_t(t.link_out.size()));
        b->max_links = std::max<int64_t>(b->max_links, int64_t(t.link_out.size()));
        for (size_t k = 0; k < t.link_out.size(); ++k) {
            lo.push_back(t.link_out[k] + pb);
            lbk.push_back(t.link_back[k] + pb);
            lq.push_back(t.link_q[k]);
        }
        geo.insert(geo.end(), t.geo.begin(), t.geo.end());
        b->pop_total += int64_t(c->Q) * c->plane;
        b->cell_total += c->plane;
        b->node_total += int64_t(c->kp.N);
    }
    b->nfill_step = int64_t(fd.size());
    fd.insert(fd.end(), fd_push.begin(), fd_push.end());
    fs.insert(fs.end(), fs_push.begin(), fs_push.end());
    HIP_TRY(hipSetDevice(device));
    const char *gv = std::getenv("IMEXLB_GRAPHS");
    b->graphs = !(gv && gv[0] == '0');
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    int cus = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    b->persist_grid = unsigned(std::max(cus, 1));
    b->nfill = int64_t(fd.size());
    const size_t pop = size_t(b->pop_total) * b->ssz, fld = size_t(b->cell_total) * sizeof(double);
    const size_t nf = std::max<size_t>(fd.size(), 1), nl = std::max<size_t>(lo.size(), 1);
    const size_t nblk = (size_t(std::max<int64_t>(b->max_links, 1)) + 255) / 256;
    imexlb_batch *B = b.get();
    int rc;
    B->fill32 = fill_fits_32(B->pop_total) && !force_slots64();
    const size_t isz = B->fill32 ? 4 : 8;
    if ((rc = balloc(B, &B->fb[0], pop)) || (rc = balloc(B, &B->fb[1], pop)) || (rc = balloc(B, &B->f, pop)) ||
        (rc = balloc(B, (void **)&B->p, fld)) || (rc = balloc(B, (void **)&B->ux, fld)) ||
        (rc = balloc(B, (void **)&B->uy, fld)) || (rc = balloc(B, (void **)&B->rv, 100 * sizeof(double))) ||
        (rc = balloc(B, (void **)&B->geo, size_t(B->cell_total))) ||
        (rc = balloc(B, &B->fill_dst, nf * isz)) || (rc = balloc(B, &B->fill_src, nf * isz)) ||
        (rc = balloc(B, (void **)&B->link_out, nl * 8)) || (rc = balloc(B, (void **)&B->link_back, nl * 8)) ||
        (rc = balloc(B, (void **)&B->link_q, nl)) ||
        (rc = balloc(B, (void **)&B->force_partial, nblk * 3 * sizeof(double))) ||
        (rc = balloc(B, (void **)&B->force_out, size_t(3 * ncase) * sizeof(double))) ||
        (rc = balloc(B, (void **)&B->mon, size_t(2 * ncase) * sizeof(unsigned long long))) ||
        (rc = balloc(B, (void **)&B->d_cases, size_t(ncase) * sizeof(BatchCase))) ||
        (rc = balloc(B, (void **)&B->queue, 64)) ||
        (rc = balloc(B, (void **)&B->push_beg, std::max<size_t>(pbeg.size(), 1) * sizeof(int32_t))) ||
        (rc = balloc(B, (void **)&B->push_dst, std::max<size_t>(pdst.size(), 1) * sizeof(int32_t))) ||
        (rc = balloc(B, (void **)&B->push_q, std::max<size_t>(pq.size(), 1))))
        return rc;
    if (!pbeg.empty()) HIP_TRY(hipMemcpy(B->push_beg, pbeg.data(), pbeg.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!pdst.empty()) HIP_TRY(hipMemcpy(B->push_dst, pdst.data(), pdst.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!pq.empty()) HIP_TRY(hipMemcpy(B->push_q, pq.data(), pq.size(), hipMemcpyHostToDevice));
    for (int i = 0; i < ncase; ++i) {
        PushTable &pt = B->cases[size_t(i)].pt;
        pt = PushTable{};
        if (beg_off[size_t(i)] < 0) continue;
        pt.beg = B->push_beg + beg_off[size_t(i)];
        pt.q = B->push_q + ent_off[size_t(i)];
        pt.dst = B->push_dst + ent_off[size_t(i)];
    }
    HIP_TRY(hipMemcpy(B->geo, geo.data(), geo.size(), hipMemcpyHostToDevice));
    HIP_TRY(upload_slots(B->fill_dst, fd, B->fill32));
    HIP_TRY(upload_slots(B->fill_src, fs, B->fill32));
    if (!lo.empty()) {
        HIP_TRY(hipMemcpy(B->link_out, lo.data(), lo.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(B->link_back, lbk.data(), lbk.size() * 8, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(B->link_q, lq.data(), lq.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(B->d_cases, B->cases.data(), size_t(ncase) * sizeof(BatchCase), hipMemcpyHostToDevice));
    if ((rc = B->build_tiles())) return rc;
    *out = b.release();
    return IMEXLB_OK;
}

int imexlb_batch_destroy(imexlb_batch *b)
{
    if (!b) return IMEXLB_OK;
    b->release();
    delete b;
    return IMEXLB_OK;
}

// InitUP (when r_vec is given; otherwise p, u hold uploaded fields), InitPDF
// and the first Collision of every case
static int batch_start(imexlb_batch *b, const double *r_vec)
{
    // both buffers start at zero: the body cells' previous values
    HIP_TRY(hipMemsetAsync(b->fb[0], 0, size_t(b->pop_total) * b->ssz, b->stream));
    HIP_TRY(hipMemsetAsync(b->fb[1], 0, size_t(b->pop_total) * b->ssz, b->stream));
    for (int i = 0; i < b->ncase; ++i) {
        StepArgs a = b->case_args(i);
        hipError_t e = r_vec ? launch_init_up(2, a, b->rv, b->stream) : hipSuccess;
        a.fout = b->pop(b->f, i);
        if (e == hipSuccess) e = launch_init_pdf(2, b->prec, a, b->stream);
        a.src = b->pop(b->f, i);
        a.dst = b->pop(b->fb[1], i);
        if (e == hipSuccess) e = launch_collide(2, b->prec, a, b->stream);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch init: ") + hipGetErrorString(e));
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->cur = 1;
    b->lazy = -1;
    b->have_pu = true;
    b->initialized = true;
    b->need_fill = true;  // the first Collision wrote no boundary copies
    return IMEXLB_OK;
}

int imexlb_batch_init(imexlb_batch *b, const double *r_vec)
{
    CHECK_BATCH(b);
    if (!r_vec) return fail(IMEXLB_E_ARG, "null r_vec");
    HIP_TRY(hipMemcpyAsync(b->rv, r_vec, 100 * sizeof(double), hipMemcpyHostToDevice, b->stream));
    return batch_start(b, r_vec);
}

int imexlb_batch_init_fields(imexlb_batch *b, const double *p, const double *u)
{
    CHECK_BATCH(b);
    if (!p || !u) return fail(IMEXLB_E_ARG, "null fields");
    HIP_TRY(hipStreamSynchronize(b->stream));
    int64_t off = 0;  // start of case i in the concatenated interior fields
    for (int i = 0; i < b->ncase; ++i) {
        const imexlb_ctx *c = b->host[size_t(i)];
        const int64_t N = c->kp.N;
        const size_t n = size_t(c->plane);
        double *dsts[3] = {b->p, b->ux, b->uy};
        for (int k = 0; k < 3; ++k) {
            std::vector<double> h(n, 0.0);
            const double *src = k == 0 ? p + off : u + 2 * off + int64_t(k - 1) * N;
            int64_t m = 0;
            for (int y = 0; y <= c->kp.ny; ++y)
                for (int x = 0; x <= c->kp.nx; ++x, ++m) h[size_t(c->pidx(x, y, 0))] = src[m];
            HIP_TRY(hipMemcpy(dsts[k] + b->cases[size_t(i)].geo_base, h.data(), n * sizeof(double),
                              hipMemcpyHostToDevice));
        }
        off += N;
    }
    return batch_start(b, nullptr);
}

int imexlb_batch_init_fields_f32(imexlb_batch *b, const float *fields, int H, int W, double p_scale,
                                 double u_scale, const double *r_vec)
{
    CHECK_BATCH(b);
    if (!fields || H <= 0 || W <= 0) return fail(IMEXLB_E_ARG, "fields");
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (r_vec) HIP_TRY(hipMemcpyAsync(b->rv, r_vec, 100 * sizeof(double), hipMemcpyHostToDevice, b->stream));
    const int64_t per = 3 * int64_t(H) * W;
    for (int i = 0; i < b->ncase; ++i) {
        hipError_t e = launch_init_given(b->case_args(i), r_vec ? b->rv : nullptr, fields + per * i, H, W, p_scale,
                                         u_scale, b->stream);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch fields: ") + hipGetErrorString(e));
    }
    return batch_start(b, nullptr);
}

int imexlb_batch_body_mask_f32(imexlb_batch *b, float *out, int H, int W, int64_t stride)
{
    CHECK_BATCH(b);
    if (!out || H <= 0 || W <= 0 || stride < int64_t(H) * W) return fail(IMEXLB_E_ARG, "mask size");
    for (int i = 0; i < b->ncase; ++i) {
        const StepArgs a = b->case_args(i);
        if (H > a.kp.ny + 1 || W > a.kp.nx + 1) return fail(IMEXLB_E_ARG, "mask larger than a case");
    }
    for (int i = 0; i < b->ncase; ++i) {
        hipError_t e = launch_body_mask(b->case_args(i), out + stride * i, H, W, b->stream);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("body mask: ") + hipGetErrorString(e));
    }
    return IMEXLB_OK;
}

int imexlb_batch_set_config(imexlb_batch *b, int block, int cells_per_thread, int persistent)
{
    CHECK_BATCH(b);
    if ((block != 64 && block != 128 && block != 256 && block != 512) ||
        (cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4))
        return fail(IMEXLB_E_ARG, "block in {64,128,256,512}, cells per thread in {1,2,4}");
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->drop_graphs();
    b->block = block;
    b->cpt = cells_per_thread;
    b->persist = persistent ? 1 : 0;
    return b->build_tiles();
}

int imexlb_batch_advance(imexlb_batch *b, int64_t nsteps)
{
    CHECK_BATCH(b);
    if (!b->initialized) return fail(IMEXLB_E_STATE, "imexlb_batch_init first");
    if (nsteps < 0) return fail(IMEXLB_E_ARG, "negative step count");
    if (b->need_fill && nsteps > 0) {  // the first step after the start, with every case's fill
        int rc = b->one_step();
        if (rc) return rc;
        --nsteps;
    }
    if (!b->graphs) {
        for (int64_t k = 0; k < nsteps; ++k) {
            int rc = b->one_step();
            if (rc) return rc;
        }
        return IMEXLB_OK;
    }
    int64_t pairs = nsteps / 2;
    for (int j = 0; j < imexlb_batch::kGraphSizes && pairs > 0; ++j) {
        const int64_t reps = pairs / imexlb_batch::kGraphPairs[j];
        if (reps <= 0) continue;
        hipGraphExec_t g;
        int rc = b->pair_graph(j, &g);
        if (rc) return rc;
        for (int64_t k = 0; k < reps; ++k) HIP_TRY(hipGraphLaunch(g, b->stream));
        pairs -= reps * imexlb_batch::kGraphPairs[j];
        // a pair returns to the same buffers; its second step streamed from 1-cur
        b->lazy = 1 - b->cur;
        b->have_pu = false;
    }
    if (nsteps % 2) {
        int rc = b->one_step();
        if (rc) return rc;
    }
    return IMEXLB_OK;
}

int imexlb_batch_tune(imexlb_batch *b, int trial_steps, int32_t best[3], double *best_ms)
{
    CHECK_BATCH(b);
    if (!b->initialized) return fail(IMEXLB_E_STATE, "imexlb_batch_init first");
    if (trial_steps < 2) trial_steps = 2;
    trial_steps += trial_steps % 2;
    struct Events {  // destroyed on every return
        hipEvent_t e[2] = {nullptr, nullptr};
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    HIP_TRY(hipEventCreate(&ev.e[0]));
    HIP_TRY(hipEventCreate(&ev.e[1]));
    double bms = 1e300;
    int bb = b->block, bc = b->cpt, bp = b->persist;
    for (int blk : {64, 128, 256, 512})
        for (int cpt : {1, 2, 4})
            for (int per : {0, 1}) {
                int rc = imexlb_batch_set_config(b, blk, cpt, per);
                if (!rc) rc = imexlb_batch_advance(b, 2);  // warm-up, builds the graph
                if (rc) return rc;
                HIP_TRY(hipEventRecord(ev.e[0], b->stream));
                if ((rc = imexlb_batch_advance(b, trial_steps))) return rc;
                HIP_TRY(hipEventRecord(ev.e[1], b->stream));
                HIP_TRY(hipEventSynchronize(ev.e[1]));
                float ms = 0.0f;
                HIP_TRY(hipEventElapsedTime(&ms, ev.e[0], ev.e[1]));
                const double per_step = double(ms) / trial_steps;
                if (per_step < bms) {
                    bms = per_step;
                    bb = blk;
                    bc = cpt;
                    bp = per;
                }
            }
    int rc = imexlb_batch_set_config(b, bb, bc, bp);
    if (rc) return rc;
    if (best) {
        best[0] = bb;
        best[1] = bc;
        best[2] = bp;
    }
    if (best_ms) *best_ms = bms;
    return IMEXLB_OK;
}

int imexlb_batch_synchronize(imexlb_batch *b)
{
    CHECK_BATCH(b);
    HIP_TRY(hipStreamSynchronize(b->stream));
    return IMEXLB_OK;
}

int imexlb_batch_forces(imexlb_batch *b, double *F)
{
    CHECK_BATCH(b);
    if (!F) return fail(IMEXLB_E_ARG, "null pointer");
    if (b->lazy < 0) return fail(IMEXLB_E_STATE, "forces need a completed step");
    for (int i = 0; i < b->ncase; ++i) {
        const int64_t lb = b->link_base[size_t(i)];
        hipError_t e = launch_forces(2, b->prec, b->fb[b->lazy], b->link_out + lb, b->link_back + lb,
                                     b->link_q + lb, b->link_count[size_t(i)], b->force_partial,
                                     b->force_out + 3 * i, b->stream);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch forces: ") + hipGetErrorString(e));
    }
    std::vector<double> h(size_t(3 * b->ncase));
    HIP_TRY(hipMemcpyAsync(h.data(), b->force_out, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::memcpy(F, h.data(), h.size() * sizeof(double));
    return IMEXLB_OK;
}

int imexlb_batch_monitor(imexlb_batch *b, double *um)
{
    CHECK_BATCH(b);
    if (!um) return fail(IMEXLB_E_ARG, "null pointer");
    if (!b->initialized) return fail(IMEXLB_E_STATE, "imexlb_batch_init first");
    for (int i = 0; i < b->ncase; ++i) {
        int rc = b->materialize(i);
        if (rc) return rc;
        hipError_t e = launch_monitor(2, b->case_args(i), b->skip[size_t(i)], b->mon + 2 * i, b->stream);
        if (e != hipSuccess) return fail(IMEXLB_E_HIP, std::string("batch monitor: ") + hipGetErrorString(e));
    }
    std::vector<unsigned long long> h(size_t(2 * b->ncase));
    HIP_TRY(hipMemcpyAsync(h.data(), b->mon, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int diverged = 0;
    for (int i = 0; i < b->ncase; ++i) {
        if (h[size_t(2 * i + 1)] != 0) {
            um[i] = std::numeric_limits<double>::quiet_NaN();
            ++diverged;
            continue;
        }
        double m;
        std::memcpy(&m, &h[size_t(2 * i)], sizeof(m));
        um[i] = std::sqrt(m);
    }
    if (diverged) return fail(IMEXLB_E_NAN, std::to_string(diverged) + " case(s) with a non-finite velocity");
    return IMEXLB_OK;
}

int imexlb_batch_download_fields(imexlb_batch *b, int i, double *p, double *u)
{
    CHECK_BATCH(b);
    if (i < 0 || i >= b->ncase || !p || !u) return fail(IMEXLB_E_ARG, "case index or null output");
    if (!b->initialized) return fail(IMEXLB_E_STATE, "imexlb_batch_init first");
    int rc = b->materialize(i);
    if (rc) return rc;
    const imexlb_ctx *c = b->host[size_t(i)];
    const size_t n = size_t(c->plane);
    std::vector<double> h(n);
    const int64_t N = c->kp.N;
    double *srcs[3] = {b->p, b->ux, b->uy};
    for (int k = 0; k < 3; ++k) {
        HIP_TRY(hipMemcpyAsync(h.data(), srcs[k] + b->cases[size_t(i)].geo_base, n * sizeof(double),
                               hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        int64_t m = 0;
        for (int y = 0; y <= c->kp.ny; ++y)
            for (int x = 0; x <= c->kp.nx; ++x, ++m) {
                const double v = h[size_t(c->pidx(x, y, 0))];
                if (k == 0)
                    p[m] = v;
                else
                    u[int64_t(k - 1) * N + m] = v;
            }
    }
    return IMEXLB_OK;
}

int imexlb_batch_info(imexlb_batch *b, int64_t out[4])
{
    if (!b || !out) return fail(IMEXLB_E_ARG, "null pointer");
    out[0] = b->node_total;
    out[1] = b->ncase;
    out[2] = int64_t(b->ntiles);
    out[3] = int64_t(b->dev_bytes);
    return IMEXLB_OK;
}

}  // extern "C"
