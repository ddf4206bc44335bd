// gppBatch -- runs the GPP problems of an instance list as one batch
// (gpp_batch_create / gpp_batch_run) on the MI355X and prints each problem's
// achtemp[iw] and asxtemp[iw] in the reference program's format
// (BGW-Kernels/gppKer.cpp:443-444), prefixed by the problem's index.  No
// counterpart in the reference, whose program runs one problem per process.
//
//   gppBatch LIST [--config C] [--repeat K] [--device D]
//
// LIST: one instance file per line (fk_instance_list_read).  Every problem
// must have the same element type (complex128: FP64; complex64: the mixed
// form).  --config 0 (default) lets the batch time its tile configurations
// and keep the fastest.
#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "fortkernels.h"

namespace {

void usage() { std::cout << "usage: gppBatch LIST [--config C] [--repeat K] [--device D]" << std::endl; }

[[noreturn]] void fail(const char *what, int rc)
{
    std::cerr << "gppBatch: " << what << " failed (" << rc << "): " << fk_last_error() << std::endl;
    std::exit(1);
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::cerr << "gppBatch: " << what << ": " << hipGetErrorString(e) << std::endl;
        std::exit(1);
    }
}

void *upload(const void *src, size_t bytes)
{
    void *d = nullptr;
    hip_check(hipMalloc(&d, bytes ? bytes : 16), "hipMalloc");
    if (bytes) hip_check(hipMemcpy(d, src, bytes, hipMemcpyHostToDevice), "hipMemcpy");
    return d;
}

}  // namespace

int main(int argc, char **argv)
{
    std::string list;
    int config = 0, repeat = 1, device = -1;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        auto next = [&]() -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        const char *v = nullptr;
        if (s == "--config") { if (!(v = next())) { usage(); return 0; } config = std::atoi(v); }
        else if (s == "--repeat") { if (!(v = next())) { usage(); return 0; } repeat = std::atoi(v); }
        else if (s == "--device") { if (!(v = next())) { usage(); return 0; } device = std::atoi(v); }
        else if (s.rfind("--", 0) == 0 || !list.empty()) { usage(); return 0; }
        else list = s;
    }
    if (list.empty() || repeat < 1) { usage(); return 0; }
    if (device >= 0) hip_check(hipSetDevice(device), "hipSetDevice");

    char **paths = nullptr;
    int32_t n = 0;
    int rc = fk_instance_list_read(list.c_str(), &paths, &n);
    if (rc != FK_OK) fail("fk_instance_list_read", rc);
    if (n == 0) { std::cerr << "gppBatch: " << list << " lists no instance" << std::endl; return 2; }

    std::vector<fk_instance *> inst(n, nullptr);
    std::vector<fk_gpp_problem> probs(n);
    std::vector<void *> dmem;
    int32_t cbytes = 0;
    for (int i = 0; i < n; ++i) {
        rc = fk_instance_open(paths[i], &inst[i]);
        if (rc != FK_OK) fail("fk_instance_open", rc);
        const fk_gpp_problem *p = fk_instance_gpp(inst[i]);
        if (!p) { std::cerr << "gppBatch: " << paths[i] << " is not a GPP instance" << std::endl; return 2; }
        const int32_t cb = fk_instance_complex_bytes(inst[i]);
        if (i == 0) cbytes = cb;
        if (cb != cbytes) { std::cerr << "gppBatch: mixed element types in " << list << std::endl; return 2; }
        // device copies of the bulk arrays; wx and the maps stay on the host
        fk_gpp_problem q = *p;
        const size_t nbc = (size_t)p->number_bands * p->ncouls * cb, gc = (size_t)p->ngpown * p->ncouls * cb;
        q.aqsmtemp = upload(p->aqsmtemp, nbc);
        q.aqsntemp = upload(p->aqsntemp, nbc);
        q.I_eps_array = upload(p->I_eps_array, gc);
        q.wtilde_array = upload(p->wtilde_array, gc);
        q.vcoul = static_cast<const double *>(upload(p->vcoul, sizeof(double) * p->ncouls));
        for (const void *d : {q.aqsmtemp, q.aqsntemp, q.I_eps_array, q.wtilde_array, (const void *)q.vcoul})
            dmem.push_back(const_cast<void *>(d));
        probs[i] = q;
        std::cout << "problem " << i << ": " << paths[i] << "  nb=" << p->number_bands << " nv=" << p->nvband
                  << " ncouls=" << p->ncouls << " ngpown=" << p->ngpown << " nfreq=" << p->nfreq << std::endl;
    }
    std::vector<const fk_gpp_problem *> pp(n);
    for (int i = 0; i < n; ++i) pp[i] = &probs[i];

    fk_options opt = fk_default_options();
    opt.memory = FK_MEM_DEVICE;
    opt.precision = cbytes == (int32_t)sizeof(fk_cfloat) ? FK_PREC_MIXED : FK_PREC_FP64;
    opt.config = config;
    fk_gpp_batch *batch = nullptr;
    rc = gpp_batch_create(n, pp.data(), &opt, &batch);
    if (rc != FK_OK) fail("gpp_batch_create", rc);

    size_t nout = 0;
    for (int i = 0; i < n; ++i) nout += 2 * (size_t)probs[i].nfreq;
    fk_cdouble *d_out = nullptr;
    hip_check(hipMalloc((void **)&d_out, sizeof(fk_cdouble) * nout), "hipMalloc outputs");
    std::vector<fk_cdouble *> ach(n), asx(n);
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        ach[i] = d_out + off;
        asx[i] = d_out + off + probs[i].nfreq;
        off += 2 * (size_t)probs[i].nfreq;
    }
    double ms_sum = 0.0;
    for (int r = 0; r < repeat; ++r) {
        float ms = 0.f;
        rc = gpp_batch_run(batch, ach.data(), asx.data(), nullptr, &ms);
        if (rc != FK_OK) fail("gpp_batch_run", rc);
        ms_sum += ms;
    }
    std::vector<fk_cdouble> h(nout);
    hip_check(hipMemcpy(h.data(), d_out, sizeof(fk_cdouble) * nout, hipMemcpyDeviceToHost), "copy");
    off = 0;
    double updates = 0.0;
    for (int i = 0; i < n; ++i) {
        const int nw = probs[i].nfreq;
        for (int iw = 0; iw < nw; ++iw)
            std::cout << "problem " << i << " achtemp[" << iw << "] = " << std::setprecision(15)
                      << std::complex<double>(h[off + iw].re, h[off + iw].im) << std::endl;
        for (int iw = 0; iw < nw; ++iw)
            std::cout << "problem " << i << " asxtemp[" << iw << "] = " << std::setprecision(15)
                      << std::complex<double>(h[off + nw + iw].re, h[off + nw + iw].im) << std::endl;
        off += 2 * (size_t)nw;
        updates += (double)probs[i].number_bands * probs[i].ngpown * probs[i].ncouls * nw;
    }
    const double ms = ms_sum / repeat;
    std::cout << std::setprecision(6) << "Batch time = " << ms << " ms per run (" << updates / (ms * 1e-3)
              << " updates/s, configuration " << gpp_batch_config(batch) << ", "
              << (opt.precision == FK_PREC_MIXED ? "mixed" : "fp64") << ", " << repeat << " run"
              << (repeat > 1 ? "s" : "") << ")" << std::endl;
    gpp_batch_destroy(batch);
    (void)hipFree(d_out);
    for (void *d : dmem) (void)hipFree(d);
    for (fk_instance *x : inst) fk_instance_close(x);
    fk_instance_list_free(paths, n);
    return 0;
}
