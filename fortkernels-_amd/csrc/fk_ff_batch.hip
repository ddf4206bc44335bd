// fk_ff_batch.hip -- many FF contractions (uneven number_bands / ncouls /
// ngpown / nFreq) in one graph launch.  No counterpart in the reference,
// whose program contracts one problem per process
// (ffkernel.deslippe.best.f90.back1:402-418, FFKer_cori/ffKernel.cpp:
// 107-131); every problem gets exactly what ff_kernel computes for it.
//
// Create: validate every problem, resolve its column map, lay out one
// device allocation for all of them (each problem's ff_plan workspace and
// output slots), upload the column maps, pick the configuration (the caller's,
// or with config 0 the fastest of the precision's configurations timed on
// this batch), then capture every problem's launches (coefficients, GEMM,
// reduction, achsDtemp) into one graph.  Run: launch the graph, then copy
// each problem's results to the caller's pointers with one scatter kernel
// (so a graph run needs no new kernel arguments).  Inputs are read at every
// run: new values in the same arrays need no new batch.
#include "fk_complex.h"
#include "fk_internal.h"

#include <algorithm>
#include <new>
#include <vector>

namespace fk {
namespace {

// Where problem i's results sit in the batch allocation (byte offsets).
struct SlotInfo {
    int64_t off_t, off_v, off_a;
    int count;                 // number_bands * nfreq
};

// Block i copies problem i's schDt / schDv / achsDtemp to the caller's
// pointers (a NULL pointer skips that output).
__global__ __launch_bounds__(256) void ff_batch_scatter_kernel(const char *base, const SlotInfo *info,
                                                               double2 *const *dst_t, double2 *const *dst_v,
                                                               double2 *const *dst_a)
{
    const int i = blockIdx.x;
    const SlotInfo in = info[i];
    const double2 *st = reinterpret_cast<const double2 *>(base + in.off_t);
    const double2 *sv = reinterpret_cast<const double2 *>(base + in.off_v);
    const double2 *sa = reinterpret_cast<const double2 *>(base + in.off_a);
    double2 *dt = dst_t[i], *dv = dst_v[i], *da = dst_a[i];
    for (int k = threadIdx.x; k < in.count; k += 256) {
        if (dt) dt[k] = st[k];
        if (dv) dv[k] = sv[k];
    }
    if (da && threadIdx.x == 0) *da = *sa;
}

struct FfBatch {
    int n = 0, config = 0;
    fk_options opt;
    std::vector<fk_ff_problem> prob;
    std::vector<std::vector<int32_t>> cols;
    std::vector<FfLayout> lay;
    std::vector<SlotInfo> slot;
    Scratch *ws = nullptr;
    void *mem = nullptr;
    size_t bytes = 0, off_info = 0, off_ptrs = 0;
    hipStream_t cs = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~FfBatch() { release(); if (cs) { (void)hipStreamSynchronize(cs); (void)hipStreamDestroy(cs); } }

    void release()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        if (cs) (void)hipStreamSynchronize(cs);
        delete ws;
        ws = nullptr;
        if (mem) (void)hipFree(mem);
        mem = nullptr;
    }

    char *at(size_t off) const { return static_cast<char *>(mem) + off; }

    // plan and allocate every problem's workspace and output slots for
    // configuration `c`, upload the column maps and the slot table
    int prepare(int c)
    {
        release();
        fk_options o = opt;
        o.config = c;
        o.kernel_ms = nullptr;
        ws = new Scratch;
        lay.assign(n, FfLayout{});
        slot.assign(n, SlotInfo{});
        for (int i = 0; i < n; ++i) {
            lay[i] = ff_plan(&prob[i], &o, *ws);
            const size_t cnt = (size_t)prob[i].number_bands * prob[i].nfreq;
            slot[i].off_t = (int64_t)ws->plan(sizeof(double2) * cnt);
            slot[i].off_v = (int64_t)ws->plan(sizeof(double2) * cnt);
            slot[i].off_a = (int64_t)ws->plan(sizeof(double2));
            slot[i].count = (int)cnt;
        }
        off_info = ws->plan(sizeof(SlotInfo) * n);
        off_ptrs = ws->plan(sizeof(void *) * 3 * n);
        bytes = ws->total();
        if (hipMalloc(&mem, bytes) != hipSuccess) {
            mem = nullptr;
            set_error("ff_batch_create: hipMalloc(%zu) failed", bytes);
            return FK_ERR_NOMEM;
        }
        FK_TRY(ws->acquire(mem, bytes, cs));
        for (int i = 0; i < n; ++i)
            FK_HIP(hipMemcpy(ws->at(lay[i].off_igp), cols[i].data(), sizeof(int32_t) * cols[i].size(),
                             hipMemcpyHostToDevice));
        FK_HIP(hipMemcpy(at(off_info), slot.data(), sizeof(SlotInfo) * n, hipMemcpyHostToDevice));
        config = c;
        return FK_OK;
    }

    // every problem's launches, in order, on stream s
    int enqueue(hipStream_t s)
    {
        fk_options o = opt;
        o.config = config;
        o.kernel_ms = nullptr;
        for (int i = 0; i < n; ++i) {
            const fk_ff_problem &p = prob[i];
            FK_TRY(ff_launch(&p, &o, lay[i], *ws, cols[i], (const double2 *)p.aqsmtemp, (const double2 *)p.aqsntemp,
                             (const double2 *)p.I_eps_array, p.vcoul, (double2 *)at(slot[i].off_t),
                             (double2 *)at(slot[i].off_v), (double2 *)at(slot[i].off_a), 1.0, s, false));
        }
        return FK_OK;
    }

    int create()
    {
        FK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        const bool mixed = opt.precision == FK_PREC_MIXED;
        if (opt.config > 0) {
            FK_TRY(prepare(opt.config));
        } else {
            // time each configuration on this batch (one warm-up run, one timed)
            const int ncfg = mixed ? ff_split_configs() : ff_fp64_configs();
            hipEvent_t e0 = nullptr, e1 = nullptr;
            FK_HIP(hipEventCreate(&e0));
            FK_HIP(hipEventCreate(&e1));
            int best = 1, rc = FK_OK;
            float best_ms = 1e30f;
            for (int c = 1; c <= ncfg && rc == FK_OK; ++c) {
                rc = prepare(c);
                if (rc == FK_OK) rc = enqueue(cs);
                if (rc == FK_OK && hipEventRecord(e0, cs) != hipSuccess) rc = FK_ERR_HIP;
                if (rc == FK_OK) rc = enqueue(cs);
                if (rc == FK_OK && hipEventRecord(e1, cs) != hipSuccess) rc = FK_ERR_HIP;
                float ms = 0.f;
                if (rc == FK_OK &&
                    (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
                    rc = FK_ERR_HIP;
                if (rc == FK_OK && ms < best_ms) { best_ms = ms; best = c; }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            if (rc != FK_OK) return rc;
            FK_TRY(prepare(best));
        }
        FK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue(cs);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cs, &g);
        graph = g;
        if (rc != FK_OK) return rc;
        if (e != hipSuccess) { set_error("ff_batch_create: graph capture: %s", hipGetErrorString(e)); return FK_ERR_HIP; }
        FK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return FK_OK;
    }

    int run(fk_cdouble *const *schDt, fk_cdouble *const *schDv, fk_cdouble *const *achs, hipStream_t s, float *ms)
    {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (ms) {
            FK_HIP(hipEventCreate(&e0));
            FK_HIP(hipEventCreate(&e1));
            FK_HIP(hipEventRecord(e0, s));
        }
        FK_HIP(hipGraphLaunch(exec, s));
        if (ms) FK_HIP(hipEventRecord(e1, s));
        std::vector<void *> ptrs(3 * (size_t)n, nullptr);
        for (int i = 0; i < n; ++i) {
            ptrs[i] = schDt[i];
            ptrs[n + i] = schDv ? schDv[i] : nullptr;
            ptrs[2 * n + i] = achs ? achs[i] : nullptr;
        }
        FK_HIP(hipMemcpyAsync(at(off_ptrs), ptrs.data(), sizeof(void *) * ptrs.size(), hipMemcpyHostToDevice, s));
        double2 *const *tab = reinterpret_cast<double2 *const *>(at(off_ptrs));
        hipLaunchKernelGGL(ff_batch_scatter_kernel, dim3(n), dim3(256), 0, s, static_cast<const char *>(mem),
                           reinterpret_cast<const SlotInfo *>(at(off_info)), tab, tab + n, tab + 2 * n);
        FK_HIP(hipGetLastError());
        if (ms) {
            FK_HIP(hipEventSynchronize(e1));
            FK_HIP(hipEventElapsedTime(ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return FK_OK;
    }
};

}  // namespace
}  // namespace fk

using namespace fk;

struct fk_ff_batch {
    FfBatch b;
};

extern "C" {

int ff_batch_create(int32_t n, const fk_ff_problem *const *problems, const fk_options *options, fk_ff_batch **batch)
{
    clear_error();
    if (!batch) { set_error("ff_batch_create: NULL batch"); return FK_ERR_ARG; }
    *batch = nullptr;
    if (n < 1 || !problems || !options) { set_error("ff_batch_create: need n >= 1 problems and options"); return FK_ERR_ARG; }
    if (options->memory != FK_MEM_DEVICE) {
        set_error("ff_batch_create: batches take device memory (FK_MEM_DEVICE)");
        return FK_ERR_UNSUPPORTED;
    }
    fk_ff_batch *h = new (std::nothrow) fk_ff_batch;
    if (!h) { set_error("ff_batch_create: out of memory"); return FK_ERR_NOMEM; }
    FfBatch &b = h->b;
    b.n = n;
    b.opt = *options;
    b.prob.resize(n);
    b.cols.resize(n);
    for (int i = 0; i < n; ++i) {
        if (!problems[i]) { delete h; set_error("ff_batch_create: NULL problem %d", i); return FK_ERR_ARG; }
        b.prob[i] = *problems[i];
        int rc = ff_check(&b.prob[i], options);
        if (rc == FK_OK) rc = resolve_columns(b.prob[i].ncouls, b.prob[i].ngpown, b.prob[i].inv_igp_index,
                                              b.prob[i].indinv, b.cols[i]);
        if (rc != FK_OK) { delete h; return rc; }
    }
    int prev = -1;
    if (options->device >= 0) {
        if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(options->device) != hipSuccess) {
            delete h;
            set_error("ff_batch_create: cannot select device %d", options->device);
            return FK_ERR_HIP;
        }
    }
    const int rc = b.create();
    if (prev >= 0) (void)hipSetDevice(prev);
    if (rc != FK_OK) { delete h; return rc; }
    *batch = h;
    return FK_OK;
}

int ff_batch_run(fk_ff_batch *batch, fk_cdouble *const *schDt, fk_cdouble *const *schDv,
                 fk_cdouble *const *achsDtemp, void *stream, float *elapsed_ms)
{
    clear_error();
    if (!batch || !schDt) { set_error("ff_batch_run: NULL argument"); return FK_ERR_ARG; }
    for (int i = 0; i < batch->b.n; ++i)
        if (!schDt[i]) { set_error("ff_batch_run: NULL schDt of problem %d", i); return FK_ERR_ARG; }
    return batch->b.run(schDt, schDv, achsDtemp, static_cast<hipStream_t>(stream), elapsed_ms);
}

int ff_batch_config(const fk_ff_batch *batch) { return batch ? batch->b.config : 0; }

size_t ff_batch_workspace_bytes(const fk_ff_batch *batch) { return batch ? batch->b.bytes : 0; }

void ff_batch_destroy(fk_ff_batch *batch) { delete batch; }

}  // extern "C"
