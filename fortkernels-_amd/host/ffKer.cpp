// ffKer -- drop-in for the reference's ffKernel program (FFKer_cori/ffKernel.cpp:302-560;
// Fortran original ffkernel.deslippe.best.f90): the static CH sum, the SX term,
// sigma 1-4 and the delta-function terms from one ff_kernel_full() call on the
// MI355X.
//
//   ffKer number_bands nvband ncouls ngpown nFreq nfreqeval [options]
//
// Without options it builds the program's inputs (aqsmtemp = aqsntemp =
// 0.5+0.5i, I_epsR = 0.5+0.5i, I_epsA = 0.5-0.5i, vcoul = 1,
// inv_igp_index[i] = i, indinv[i] = i, ekq[n1] = -10 + n1, dFreqGrid[f] = 2f,
// pref, dFreqBrd = 0.1i, freqevalmin = 0, freqevalstep = 0.5, pref_zb =
// 0.5/3.14; ffkernel.deslippe.best.f90:111-206, FFKer_cori/ffKernel.cpp:343-450)
// in host memory and prints, as the C++ program does ("( %f, %f) ",
// Complex.h:91-94, :528-533), achsDtemp, asxDtemp[0] and achDtemp_cor[0]; then
// the Fortran program's "Answer:" line (:660) and every output at full
// precision.  A wrong argument count, or ngpown > ncouls, prints the
// reference's message and exits with status 0 (:305-325).
//
// The values follow the Fortran program by default.  The C++ program departs
// from it in asxDtemp and achDtemp_cor (its evaluation axis, its aliased
// sigma-1 output and its never-computed schDt_matrix); --semantics 2 prints
// the C++ program's values instead.  achsDtemp: the C++ program reads
// frequency index 1 (:122), the Fortran the first frequency (:263); the
// default follows the program this replaces; --achs-freq 0 gives the
// Fortran.
//
// Options:
//   --random SEED    seeded inputs generated on the device: aqs*, I_epsR,
//                    I_epsA ~ U(-0.5,0.5) + iU(-0.5,0.5), vcoul U(0.1,10)
//   --achs-freq F    frequency index of achsDtemp (default 1)
//   --semantics N    0: the Fortran program (default); 1: the C++ port's
//                    arithmetic on the contracted schDt; 2: the C++ program
//                    as it runs, schDt_matrix zero (fk_ff_semantics)
//   --mixed          both contractions in split BF16 (FP64 evaluation)
//   --contraction    only the contraction schDt (ff_kernel), as before the
//                    evaluation existed; with --both also for I_epsA
//   --repeat K       run K times, print the mean device time
//   --device D       HIP device ordinal
//   --dump FILE      write the inputs of this run (the evaluation's) to an
//                    instance file
//   --input FILE     run the FF problem of an instance file (its own axes,
//                    grid, weights and maps) instead of the six numbers
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "fortkernels.h"

namespace {

void usage_params()
{
    std::cout << "./a.out number_bands nvband ncouls ngpown nFreq nfreqeval "
                 "[--random SEED] [--achs-freq F] [--semantics N] [--mixed] [--contraction [--both]] [--repeat K] "
                 "[--device D] [--dump FILE]\n./a.out --input FILE [--semantics N] [--mixed] [--repeat K] [--device D]"
              << std::endl;
}

[[noreturn]] void fail(const char *what, int rc)
{
    std::cerr << "ffKer: " << what << " failed (" << rc << "): " << fk_last_error() << std::endl;
    std::exit(1);
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::cerr << "ffKer: " << what << ": " << hipGetErrorString(e) << std::endl;
        std::exit(1);
    }
}

struct Args {
    int nb = 0, nv = 0, ncouls = 0, ngpown = 0, nfreq = 0, nfreqeval = 0;
    bool random = false, both = false, mixed = false, contraction = false;
    std::string input, dump;
    unsigned long long seed = 0;
    int achs_freq = 1, repeat = 1, device = -1, semantics = 0;
};

bool parse(int argc, char **argv, Args &a)
{
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        auto next = [&](void) -> const char * { return i + 1 < argc ? argv[++i] : nullptr; };
        if (s.rfind("--", 0) != 0) { pos.push_back(s); continue; }
        const char *v = nullptr;
        if (s == "--random") { if (!(v = next())) return false; a.random = true; a.seed = std::strtoull(v, nullptr, 10); }
        else if (s == "--achs-freq") { if (!(v = next())) return false; a.achs_freq = std::atoi(v); }
        else if (s == "--semantics") { if (!(v = next())) return false; a.semantics = std::atoi(v); }
        else if (s == "--both") a.both = true;
        else if (s == "--mixed") a.mixed = true;
        else if (s == "--contraction") a.contraction = true;
        else if (s == "--repeat") { if (!(v = next())) return false; a.repeat = std::atoi(v); }
        else if (s == "--device") { if (!(v = next())) return false; a.device = std::atoi(v); }
        else if (s == "--input") { if (!(v = next())) return false; a.input = v; }
        else if (s == "--dump") { if (!(v = next())) return false; a.dump = v; }
        else return false;
    }
    if (!a.input.empty()) return pos.empty() && a.repeat >= 1 && !a.random && !a.contraction && a.dump.empty();
    if (pos.size() != 6) return false;
    int *dst[] = {&a.nb, &a.nv, &a.ncouls, &a.ngpown, &a.nfreq, &a.nfreqeval};
    for (int k = 0; k < 6; ++k) *dst[k] = std::atoi(pos[k].c_str());
    return a.repeat >= 1;
}

void print_c(const char *name, fk_cdouble z) { std::printf("%s = ( %f, %f) \n", name, z.re, z.im); }

// One ff_kernel_full call (or K), then the program's output lines.
void run_full_and_print(const fk_ff_full_problem &p, fk_options opt, const Args &a)
{
    const int nb = p.number_bands, C = p.ncouls, G = p.ngpown, nf = p.nfreq, ne = p.nfreqeval;
    const bool on_dev = opt.memory == FK_MEM_DEVICE;
    float kms = 0.f;
    opt.kernel_ms = &kms;
    // outputs: achs, then five [ne] arrays
    const size_t nout = 1 + 5 * (size_t)ne;
    std::vector<fk_cdouble> h(nout);
    fk_cdouble *d_out = nullptr;
    fk_cdouble *o = h.data();
    if (on_dev) {
        hip_check(hipMalloc((void **)&d_out, sizeof(fk_cdouble) * nout), "hipMalloc outputs");
        o = d_out;
    }
    fk_ff_full_outputs out;
    out.achsDtemp = o; out.asxDtemp = o + 1; out.achDtemp = o + 1 + ne; out.achDtemp_cor = o + 1 + 2 * ne;
    out.achDtemp_corb = o + 1 + 3 * ne; out.ach2Dtemp = o + 1 + 4 * ne; out.schDt = nullptr;
    double ms_sum = 0.0;
    const auto t_k = std::chrono::high_resolution_clock::now();
    for (int r = 0; r < a.repeat; ++r) {
        const int rc = ff_kernel_full(&p, &opt, &out);
        if (rc != FK_OK) fail("ff_kernel_full", rc);
        ms_sum += kms;
    }
    const std::chrono::duration<double> kernel_time = std::chrono::high_resolution_clock::now() - t_k;
    if (on_dev) {
        hip_check(hipMemcpy(h.data(), d_out, sizeof(fk_cdouble) * nout, hipMemcpyDeviceToHost), "copy");
        (void)hipFree(d_out);
    }
    const fk_cdouble *asx = &h[1], *ach = asx + ne, *cor = ach + ne, *corb = cor + ne, *a2 = corb + ne;
    print_c("achsDtemp", h[0]);
    print_c("asxDtemp", asx[0]);
    print_c("achDtemp_cor", cor[0]);
    std::printf(" Answer: (%.15e,%.15e) (%.15e,%.15e) (%.15e,%.15e)\n", cor[0].re, cor[0].im, ach[0].re,
                ach[0].im, asx[0].re, asx[0].im);
    std::printf("achsDtemp (full precision) = (%.15e, %.15e)\n", h[0].re, h[0].im);
    const char *names[5] = {"asxDtemp", "achDtemp", "achDtemp_cor", "achDtemp_corb", "ach2Dtemp"};
    const fk_cdouble *vals[5] = {asx, ach, cor, corb, a2};
    for (int k = 0; k < 5; ++k)
        for (int iw = 0; iw < ne; ++iw)
            std::printf("%s[%d] = (%.15e,%.15e)\n", names[k], iw, vals[k][iw].re, vals[k][iw].im);
    const double flops = 8.0 * (double)G * nf * nb * C;
    const double ms = ms_sum / a.repeat;
    std::printf("Device time = %.4f ms per call (both contractions and the evaluation; the I_epsR "
                "contraction alone is %.3e flops, %.2f TFLOP/s over the whole call; %d call%s)\n",
                ms, flops, flops / (ms * 1e-3) / 1e12, a.repeat, a.repeat > 1 ? "s" : "");
    std::cout << "********** Total Kernel Time **********= " << kernel_time.count() << " secs" << std::endl;
}

// The inputs of this run as an instance file (device arrays copied back).
void dump_instance(const std::string &path, const fk_ff_full_problem &p, bool on_dev)
{
    fk_ff_full_problem q = p;
    const size_t nbc = (size_t)p.number_bands * p.ncouls, xc = (size_t)p.nfreq * p.ngpown * p.ncouls;
    std::vector<fk_cdouble> m, n, r, x;
    std::vector<double> v;
    if (on_dev) {
        m.resize(nbc); n.resize(nbc); r.resize(xc); x.resize(xc); v.resize(p.ncouls);
        hip_check(hipMemcpy(m.data(), p.aqsmtemp, nbc * sizeof(fk_cdouble), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(n.data(), p.aqsntemp, nbc * sizeof(fk_cdouble), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(r.data(), p.I_epsR_array, xc * sizeof(fk_cdouble), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(x.data(), p.I_epsA_array, xc * sizeof(fk_cdouble), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(v.data(), p.vcoul, sizeof(double) * v.size(), hipMemcpyDeviceToHost), "dump copy");
        q.aqsmtemp = m.data(); q.aqsntemp = n.data(); q.I_epsR_array = r.data(); q.I_epsA_array = x.data();
        q.vcoul = v.data();
    }
    const int rc = fk_ff_instance_write(path.c_str(), &q);
    if (rc != FK_OK) fail("fk_ff_instance_write", rc);
    std::cout << "Inputs written to " << path << std::endl;
}

// --input: the problem of an instance file, host memory.
int run_input(const Args &a)
{
    fk_instance *inst = nullptr;
    int rc = fk_instance_open(a.input.c_str(), &inst);
    if (rc != FK_OK) fail("fk_instance_open", rc);
    const fk_ff_full_problem *p = fk_instance_ff(inst);
    if (!p) {
        std::cerr << "ffKer: " << a.input << " holds a GPP problem, not an FF one" << std::endl;
        fk_instance_close(inst);
        return 2;
    }
    if (a.device >= 0) hip_check(hipSetDevice(a.device), "hipSetDevice");
    std::cout << "Instance = " << a.input << "\n number_bands = " << p->number_bands << "\n nvband = " << p->nvband
              << "\n ncouls = " << p->ncouls << "\n ngpown = " << p->ngpown << "\n nFreq = " << p->nfreq
              << "\n nfreqeval = " << p->nfreqeval << std::endl;
    fk_options opt = fk_default_options();
    opt.memory = FK_MEM_HOST;
    opt.precision = a.mixed ? FK_PREC_MIXED : FK_PREC_FP64;
    opt.device = a.device;
    fk_ff_full_problem q = *p;
    q.semantics = a.semantics;
    run_full_and_print(q, opt, a);
    fk_instance_close(inst);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Args a;
    const bool parsed = parse(argc, argv, a);
    if (parsed && !a.input.empty()) return run_input(a);
    if (!parsed) {
        std::cout << "Incorrect Parameters!!! The correct form is " << std::endl;
        usage_params();
        return 0;
    }
    if (a.ngpown > a.ncouls) {
        std::cout << "Incorrect Parameters!!! ngpown cannot be greater than ncouls. The correct form is " << std::endl;
        usage_params();
        return 0;
    }
    const int nb = a.nb, C = a.ncouls, G = a.ngpown, nf = a.nfreq, ne = a.nfreqeval;
    if (nb < 1 || C < 1 || G < 1 || nf < 1 || a.achs_freq < 0 || a.achs_freq >= nf) {
        std::cerr << "ffKer: need number_bands, ncouls, ngpown, nFreq >= 1 and 0 <= achs-freq < nFreq" << std::endl;
        return 2;
    }
    if (!a.contraction && (nf < 2 || ne < 1 || a.nv < 0 || a.nv > nb)) {
        std::cerr << "ffKer: the evaluation needs nFreq >= 2, nfreqeval >= 1 and 0 <= nvband <= number_bands"
                  << std::endl;
        return 2;
    }
    const auto t_start = std::chrono::high_resolution_clock::now();
    if (a.device >= 0) hip_check(hipSetDevice(a.device), "hipSetDevice");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::cout << "Device = " << prop.name << " (" << fk_build_target() << ", " << prop.multiProcessorCount
              << " CUs)" << "\n number_bands = " << nb << "\n nvband = " << a.nv << "\n ncouls = " << C
              << "\n ngpown = " << G << "\n nFreq = " << nf << "\n nfreqeval = " << ne << std::endl;

    std::vector<int32_t> inv(G), indinv(C + 1);
    for (int i = 0; i < G; ++i) inv[i] = i;
    for (int i = 0; i < C; ++i) indinv[i] = i;
    indinv[C] = C - 1;
    // the program's axes, ffkernel.deslippe.best.f90:124-171
    std::vector<double> ekq(nb), grid(nf), pref(nf);
    std::vector<fk_cdouble> brd(nf, fk_cdouble{0.0, 0.1});
    for (int n1 = 0; n1 < nb; ++n1) ekq[n1] = -10.0 + n1;
    for (int f = 0; f < nf; ++f) grid[f] = 2.0 * f;
    for (int f = 0; f < nf; ++f) pref[f] = f < nf - 1 ? (grid[f + 1] - grid[f]) / 3.14 : (nf > 1 ? pref[f - 1] : 0.0);
    pref[0] *= 0.5;
    pref[nf - 1] *= 0.5;

    const size_t nbc = (size_t)nb * C, xc = (size_t)nf * G * C;
    const int narr = (!a.contraction || a.both) ? 2 : 1;
    std::cout << "Memory Used = " << (double)(2 * nbc + narr * xc) * sizeof(fk_cdouble) / (1024.0 * 1024 * 1024)
              << " GB" << std::endl;

    fk_options opt = fk_default_options();
    opt.device = a.device;
    opt.precision = a.mixed ? FK_PREC_MIXED : FK_PREC_FP64;
    std::vector<fk_cdouble> h_aqs, h_x[2];
    std::vector<double> h_v;
    void *d_m = nullptr, *d_n = nullptr, *d_x[2] = {nullptr, nullptr};
    double *d_v = nullptr;
    const void *pm, *pn, *px[2] = {nullptr, nullptr};
    const double *pv;
    if (!a.random) {
        opt.memory = FK_MEM_HOST;
        h_aqs.assign(nbc, fk_cdouble{0.5, 0.5});
        h_x[0].assign(xc, fk_cdouble{0.5, 0.5});                      // I_epsR
        if (narr == 2) h_x[1].assign(xc, fk_cdouble{0.5, -0.5});      // I_epsA
        h_v.assign(C, 1.0);
        pm = pn = h_aqs.data();
        px[0] = h_x[0].data();
        px[1] = narr == 2 ? h_x[1].data() : nullptr;
        pv = h_v.data();
    } else {
        opt.memory = FK_MEM_DEVICE;
        hip_check(hipMalloc(&d_m, nbc * sizeof(fk_cdouble)), "hipMalloc aqsmtemp");
        hip_check(hipMalloc(&d_n, nbc * sizeof(fk_cdouble)), "hipMalloc aqsntemp");
        for (int k = 0; k < narr; ++k) hip_check(hipMalloc(&d_x[k], xc * sizeof(fk_cdouble)), "hipMalloc I_eps");
        hip_check(hipMalloc((void **)&d_v, C * sizeof(double)), "hipMalloc vcoul");
        int rc = fk_fill_complex(d_m, (int64_t)nbc, a.seed, 1, -0.5, 0.5, -0.5, 0.5, 0, nullptr);
        if (rc == FK_OK) rc = fk_fill_complex(d_n, (int64_t)nbc, a.seed, 2, -0.5, 0.5, -0.5, 0.5, 0, nullptr);
        for (int k = 0; k < narr && rc == FK_OK; ++k)
            rc = fk_fill_complex(d_x[k], (int64_t)xc, a.seed, 6 + k, -0.5, 0.5, -0.5, 0.5, 0, nullptr);
        if (rc == FK_OK) rc = fk_fill_real(d_v, C, a.seed, 5, 0.1, 10.0, nullptr);
        if (rc != FK_OK) fail("input fill", rc);
        hip_check(hipDeviceSynchronize(), "input fill");
        pm = d_m; pn = d_n; px[0] = d_x[0]; px[1] = d_x[1]; pv = d_v;
    }
    const std::chrono::duration<double> preloop = std::chrono::high_resolution_clock::now() - t_start;
    std::cout << "********** PreLoop **********= " << preloop.count() << " secs" << std::endl;
    std::cout << "starting Kernels" << std::endl;
    float kms = 0.f;
    opt.kernel_ms = &kms;

    if (!a.contraction) {
        fk_ff_full_problem p;
        std::memset(&p, 0, sizeof(p));
        p.number_bands = nb; p.nvband = a.nv; p.ncouls = C; p.ngpown = G; p.nfreq = nf; p.nfreqeval = ne;
        p.aqsmtemp = pm; p.aqsntemp = pn; p.I_epsR_array = px[0]; p.I_epsA_array = px[1]; p.vcoul = pv;
        p.inv_igp_index = inv.data(); p.indinv = indinv.data();
        p.ekq = ekq.data(); p.dFreqGrid = grid.data(); p.pref = pref.data(); p.dFreqBrd = brd.data();
        p.pref_zb = 0.5 / 3.14; p.freqevalmin = 0.0; p.freqevalstep = 0.5; p.occ = 1.0;
        p.achs_freq = a.achs_freq;
        p.semantics = a.semantics;
        if (!a.dump.empty()) dump_instance(a.dump, p, a.random);
        run_full_and_print(p, opt, a);
    } else {
        const size_t nout = (size_t)nb * nf;
        std::vector<fk_cdouble> schDt[2], achs(2);
        fk_cdouble *d_out = nullptr;
        if (a.random) hip_check(hipMalloc((void **)&d_out, sizeof(fk_cdouble) * 2 * (nout + 1)), "hipMalloc outputs");
        double kms_sum[2] = {0.0, 0.0};
        const auto t_k = std::chrono::high_resolution_clock::now();
        for (int k = 0; k < narr; ++k) {
            schDt[k].resize(nout);
            fk_ff_problem p;
            std::memset(&p, 0, sizeof(p));
            p.number_bands = nb; p.ncouls = C; p.ngpown = G; p.nfreq = nf;
            p.aqsmtemp = pm; p.aqsntemp = pn; p.I_eps_array = px[k]; p.vcoul = pv;
            p.inv_igp_index = inv.data(); p.indinv = indinv.data();
            p.achs_freq = a.achs_freq;
            fk_cdouble *o_t = a.random ? d_out + k * (nout + 1) : schDt[k].data();
            fk_cdouble *o_a = a.random ? d_out + k * (nout + 1) + nout : &achs[k];
            for (int r = 0; r < a.repeat; ++r) {
                const int rc = ff_kernel(&p, &opt, o_t, nullptr, o_a);
                if (rc != FK_OK) fail("ff_kernel", rc);
                kms_sum[k] += kms;
            }
            if (a.random) {
                hip_check(hipMemcpy(schDt[k].data(), o_t, sizeof(fk_cdouble) * nout, hipMemcpyDeviceToHost), "copy");
                hip_check(hipMemcpy(&achs[k], o_a, sizeof(fk_cdouble), hipMemcpyDeviceToHost), "copy");
            }
        }
        const std::chrono::duration<double> kernel_time = std::chrono::high_resolution_clock::now() - t_k;
        if (d_out) (void)hipFree(d_out);
        print_c("achsDtemp", achs[0]);
        std::printf("achsDtemp (full precision) = (%.15e, %.15e)\n", achs[0].re, achs[0].im);
        const char *names[2] = {"I_epsR", "I_epsA"};
        const double flops = 8.0 * (double)G * nf * nb * C;
        for (int k = 0; k < narr; ++k) {
            double sr = 0, si = 0;
            for (size_t i = 0; i < nout; ++i) { sr += schDt[k][i].re; si += schDt[k][i].im; }
            std::printf("schDt[%s]: sum over %d bands x %d frequencies = (%.15e, %.15e); schDt[0][0] = (%.15e, %.15e)\n",
                        names[k], nb, nf, sr, si, schDt[k][0].re, schDt[k][0].im);
            const double ms = kms_sum[k] / a.repeat;
            std::printf("Kernel time [%s] = %.4f ms per call (%.2f TFLOP/s, %d call%s)\n", names[k], ms,
                        flops / (ms * 1e-3) / 1e12, a.repeat, a.repeat > 1 ? "s" : "");
        }
        std::cout << "********** Total Kernel Time **********= " << kernel_time.count() << " secs" << std::endl;
    }
    for (void *ptr : {d_m, d_n, d_x[0], d_x[1], (void *)d_v})
        if (ptr) (void)hipFree(ptr);
    return 0;
}
