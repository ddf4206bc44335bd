// gppKer -- drop-in for the reference's gppKer program (BGW-Kernels/gppKer.cpp:161-450)
// with the band loop replaced by one gpp_kernel() call on the MI355X.
//
//   gppKer <number_bands> <number_valence_bands> <number_plane_waves> <nodes_per_mpi_group> [options]
//   gppKer <number_bands> <number_valence_bands> <number_plane_waves> <nodes_per_mpi_group> 1 [options]
//
// The five-number form is the command line of the sequential program
// gppKerSeq.cpp (:264-276) with gppsum = 1: the same inputs, its band loop
// (FK_GPP_GPPSUM semantics, FP64) and the same output lines.  Only gppsum = 1
// is built; the four-number form runs BGW-Kernels/gppKer.cpp's loop.
//
// Without options it builds the reference's inputs (every aqsmtemp, aqsntemp,
// I_eps_array and wtilde_array entry 0.5+0.5i, vcoul 1, inv_igp_index[i] =
// (i+1)*ncouls/ngpown, indinv[i] = i; :275-299) in host memory, hands them
// over as FK_MEM_HOST and prints achtemp[iw] at precision 15 (:443-444),
// asxtemp[iw] and the time taken (:446).  A wrong argument count prints the
// usage and exits with status 0, as the reference does (:164-168).
//
// Options (all after the four numbers):
//   --random SEED   seeded synthetic inputs generated on the device:
//                   aqs*, I_eps ~ U(-0.5,0.5) + iU(-0.5,0.5), wtilde re
//                   U(0.5,2.0) im U(-0.01,0.01), vcoul U(0.1,10),
//                   inv_igp_index[i] = i*ncouls/ngpown, indinv[ncouls] = ncouls-1
//   --near-pole     with --random: wtilde re ~ U(3.0,5.0)
//   --nfreq N       number of frequencies (default 3; wx[iw] = 3 + iw)
//   --mixed         complex64 storage, FK_PREC_MIXED: the band contraction on
//                   int8 MFMA by the exact-accumulation split (config 0)
//   --unmasked      the GPU variants' unmasked sum (gppKer_gpuComplex.cpp:171-178)
//   --acht-n1       also compute and print acht_n1_loc[n1] (:424)
//   --achstemp      also compute and print the static-CH remainder achstemp of
//                   the Fortran program (gppkernel.deslippe.09.30.2014.f90:
//                   192-230; gpp_kernel_ex)
//   --repeat K      run the kernel K times and print the mean kernel time
//   --device D      HIP device ordinal
//   --dump FILE     write the inputs of this run to an instance file
//   --input FILE    run the problem of an instance file (fk_instance_open)
//                   instead of the four numbers: sizes, wx, column maps and
//                   arrays come from the file; complex64 files run mixed
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "fortkernels.h"

namespace {

void usage()
{
    std::cout << "The correct form of input is : " << std::endl;
    std::cout << " ./a.out <number_bands> <number_valence_bands> <number_plane_waves> <nodes_per_mpi_group> "
                 "[gppsum=1] [--random SEED] [--near-pole] [--nfreq N] [--mixed] [--unmasked] [--acht-n1] [--achstemp] [--repeat K] "
                 "[--device D] [--dump FILE]\n ./a.out --input FILE [--unmasked] [--acht-n1] [--achstemp] [--repeat K] "
                 "[--device D]"
              << std::endl;
}

[[noreturn]] void fail(const char *what, int rc)
{
    std::cerr << "gppKer: " << what << " failed (" << rc << "): " << fk_last_error() << std::endl;
    std::exit(1);
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::cerr << "gppKer: " << what << ": " << hipGetErrorString(e) << std::endl;
        std::exit(1);
    }
}

struct Args {
    int nb = 0, nv = 0, ncouls = 0, npg = 0, gppsum = -1;   // -1: the four-number form
    bool random = false, near_pole = false, mixed = false, unmasked = false, acht_n1 = false, achstemp = false;
    unsigned long long seed = 0;
    int nfreq = 3, repeat = 1, device = -1;
    std::string input, dump;
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
        else if (s == "--near-pole") a.near_pole = true;
        else if (s == "--nfreq") { if (!(v = next())) return false; a.nfreq = std::atoi(v); }
        else if (s == "--mixed") a.mixed = true;
        else if (s == "--unmasked") a.unmasked = true;
        else if (s == "--acht-n1") a.acht_n1 = true;
        else if (s == "--achstemp") a.achstemp = true;
        else if (s == "--repeat") { if (!(v = next())) return false; a.repeat = std::atoi(v); }
        else if (s == "--device") { if (!(v = next())) return false; a.device = std::atoi(v); }
        else if (s == "--input") { if (!(v = next())) return false; a.input = v; }
        else if (s == "--dump") { if (!(v = next())) return false; a.dump = v; }
        else return false;
    }
    if (!a.input.empty()) return pos.empty() && a.repeat >= 1 && !a.random && a.dump.empty();
    if (pos.size() != 4 && pos.size() != 5) return false;
    if (pos.size() == 5) a.gppsum = std::atoi(pos[4].c_str());
    a.nb = std::atoi(pos[0].c_str());
    a.nv = std::atoi(pos[1].c_str());
    a.ncouls = std::atoi(pos[2].c_str());
    a.npg = std::atoi(pos[3].c_str());
    return a.repeat >= 1;
}

// One gpp_kernel_ex call (or K), then the reference's output lines.
void run_and_print(const fk_gpp_problem &p, fk_options opt, const Args &a)
{
    const int nb = p.number_bands, nfreq = p.nfreq, ngpown = p.ngpown, ncouls = p.ncouls;
    const bool on_dev = opt.memory == FK_MEM_DEVICE;
    // outputs: host for FK_MEM_HOST, device (copied back) for FK_MEM_DEVICE
    std::vector<fk_cdouble> achtemp(nfreq), asxtemp(nfreq), acht_n1(a.acht_n1 ? nb : 0), achstemp(1);
    fk_cdouble *o_ach = achtemp.data(), *o_asx = asxtemp.data(), *o_n1 = a.acht_n1 ? acht_n1.data() : nullptr;
    fk_cdouble *o_achs = a.achstemp ? achstemp.data() : nullptr;
    fk_cdouble *d_out = nullptr;
    if (on_dev) {
        hip_check(hipMalloc((void **)&d_out, sizeof(fk_cdouble) * (2 * nfreq + nb + 1)), "hipMalloc outputs");
        o_ach = d_out;
        o_asx = d_out + nfreq;
        o_n1 = a.acht_n1 ? d_out + 2 * nfreq : nullptr;
        o_achs = a.achstemp ? d_out + 2 * nfreq + nb : nullptr;
    }
    fk_gpp_outputs outs;
    outs.achtemp = o_ach; outs.asxtemp = o_asx; outs.acht_n1_loc = o_n1; outs.achstemp = o_achs;
    float kms = 0.f;
    double kms_sum = 0.0;
    opt.kernel_ms = &kms;
    const auto t0 = std::chrono::high_resolution_clock::now();
    for (int r = 0; r < a.repeat; ++r) {
        int rc = gpp_kernel_ex(&p, &opt, &outs);
        if (rc != FK_OK) fail("gpp_kernel", rc);
        kms_sum += kms;
    }
    if (on_dev) hip_check(hipDeviceSynchronize(), "gpp_kernel");
    const std::chrono::duration<double> elapsed = std::chrono::high_resolution_clock::now() - t0;
    if (on_dev) {
        hip_check(hipMemcpy(achtemp.data(), o_ach, sizeof(fk_cdouble) * nfreq, hipMemcpyDeviceToHost), "copy");
        hip_check(hipMemcpy(asxtemp.data(), o_asx, sizeof(fk_cdouble) * nfreq, hipMemcpyDeviceToHost), "copy");
        if (a.acht_n1)
            hip_check(hipMemcpy(acht_n1.data(), o_n1, sizeof(fk_cdouble) * nb, hipMemcpyDeviceToHost), "copy");
        if (a.achstemp)
            hip_check(hipMemcpy(achstemp.data(), o_achs, sizeof(fk_cdouble), hipMemcpyDeviceToHost), "copy");
    }

    for (int iw = 0; iw < nfreq; ++iw)
        std::cout << "achtemp[" << iw << "] = " << std::setprecision(15)
                  << std::complex<double>(achtemp[iw].re, achtemp[iw].im) << std::endl;
    for (int iw = 0; iw < nfreq; ++iw)
        std::cout << "asxtemp[" << iw << "] = " << std::setprecision(15)
                  << std::complex<double>(asxtemp[iw].re, asxtemp[iw].im) << std::endl;
    if (a.acht_n1)
        for (int n1 = 0; n1 < nb; ++n1)
            std::cout << "acht_n1_loc[" << n1 << "] = " << std::setprecision(15)
                      << std::complex<double>(acht_n1[n1].re, acht_n1[n1].im) << std::endl;
    if (a.achstemp)
        std::cout << "achstemp = " << std::setprecision(15) << std::complex<double>(achstemp[0].re, achstemp[0].im)
                  << std::endl;
    const double updates = (double)nb * ngpown * ncouls * nfreq;
    std::cout << std::setprecision(6) << "Kernel time = " << kms_sum / a.repeat << " ms per call ("
              << updates / (kms_sum / a.repeat * 1e-3) << " updates/s, " << a.repeat << " call"
              << (a.repeat > 1 ? "s" : "") << ", " << (opt.precision == FK_PREC_MIXED ? "mixed" : "fp64") << ")"
              << std::endl;
    std::cout << "********** Chrono Time Taken **********= " << elapsed.count() << " secs" << std::endl;

    if (d_out) (void)hipFree(d_out);
}

// The inputs of this run as an instance file (device arrays copied back).
void dump_instance(const std::string &path, const fk_gpp_problem &p, bool c64, bool on_dev)
{
    fk_gpp_problem q = p;
    const size_t esz = c64 ? sizeof(fk_cfloat) : sizeof(fk_cdouble);
    const size_t nbc = (size_t)p.number_bands * p.ncouls, gc = (size_t)p.ngpown * p.ncouls;
    std::vector<char> m, n, e, w;
    std::vector<double> v;
    if (on_dev) {
        m.resize(nbc * esz); n.resize(nbc * esz); e.resize(gc * esz); w.resize(gc * esz); v.resize(p.ncouls);
        hip_check(hipMemcpy(m.data(), p.aqsmtemp, m.size(), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(n.data(), p.aqsntemp, n.size(), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(e.data(), p.I_eps_array, e.size(), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(w.data(), p.wtilde_array, w.size(), hipMemcpyDeviceToHost), "dump copy");
        hip_check(hipMemcpy(v.data(), p.vcoul, sizeof(double) * v.size(), hipMemcpyDeviceToHost), "dump copy");
        q.aqsmtemp = m.data(); q.aqsntemp = n.data(); q.I_eps_array = e.data(); q.wtilde_array = w.data();
        q.vcoul = v.data();
    }
    const int rc = fk_gpp_instance_write(path.c_str(), &q, c64 ? 1 : 0);
    if (rc != FK_OK) fail("fk_gpp_instance_write", rc);
    std::cout << "Inputs written to " << path << std::endl;
}

// --input: the problem of an instance file, host memory.
int run_input(const Args &a)
{
    fk_instance *inst = nullptr;
    int rc = fk_instance_open(a.input.c_str(), &inst);
    if (rc != FK_OK) fail("fk_instance_open", rc);
    const fk_gpp_problem *p = fk_instance_gpp(inst);
    if (!p) {
        std::cerr << "gppKer: " << a.input << " holds an FF problem, not a GPP one" << std::endl;
        fk_instance_close(inst);
        return 2;
    }
    if (a.device >= 0) hip_check(hipSetDevice(a.device), "hipSetDevice");
    const bool c64 = fk_instance_complex_bytes(inst) == (int32_t)sizeof(fk_cfloat);
    std::cout << "Instance = " << a.input << " (" << (c64 ? "complex64" : "complex128") << ")" << std::endl;
    std::cout << "number_bands = " << p->number_bands << "\t nvband = " << p->nvband << "\t ncouls = " << p->ncouls
              << "\t ngpown = " << p->ngpown << "\t nend = " << p->nfreq << "\t limitone = " << p->limitone
              << "\t limittwo = " << p->limittwo << std::endl;
    fk_options opt = fk_default_options();
    opt.memory = FK_MEM_HOST;
    opt.precision = c64 ? FK_PREC_MIXED : FK_PREC_FP64;
    opt.mask = a.unmasked ? 0 : 1;
    opt.device = a.device;
    run_and_print(*p, opt, a);
    fk_instance_close(inst);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Args a;
    if (!parse(argc, argv, a)) {
        usage();
        return 0;
    }
    if (!a.input.empty()) return run_input(a);
    if (a.gppsum >= 0 && (a.gppsum != 1 || a.mixed || a.unmasked || a.acht_n1 || a.achstemp)) {
        std::cerr << "gppKer: the five-number form (gppKerSeq.cpp) is built for gppsum = 1, in FP64, without "
                     "--mixed, --unmasked, --acht-n1 or --achstemp; the four-number form runs "
                     "BGW-Kernels/gppKer.cpp's loop" << std::endl;
        return 2;
    }
    if (a.nb < 1 || a.nv < 0 || a.nv > a.nb || a.ncouls < 1 || a.npg < 1 || a.npg > a.ncouls) {
        std::cerr << "gppKer: need number_bands >= 1, 0 <= number_valence_bands <= number_bands, "
                     "1 <= nodes_per_mpi_group <= number_plane_waves" << std::endl;
        return 2;
    }
    const int nb = a.nb, nv = a.nv, ncouls = a.ncouls, npg = a.npg, nfreq = a.nfreq;
    const int ngpown = ncouls / npg;   // npes = 1 (:175-176)
    const double e_lk = 10, dw = 1, e_n1kq = 6.0, to1 = 1e-6, gamma = 0.5, sexcut = 4.0;
    const double limitone = 1.0 / (to1 * 4.0), limittwo = 0.5 * 0.5;

    if (a.device >= 0) hip_check(hipSetDevice(a.device), "hipSetDevice");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    std::cout << "Device = " << prop.name << " (" << fk_build_target() << ", " << prop.multiProcessorCount
              << " CUs)" << std::endl;
    std::cout << "number_bands = " << nb << "\t nvband = " << nv << "\t ncouls = " << ncouls
              << "\t nodes_per_group  = " << npg << "\t ngpown = " << ngpown << "\t nend = " << nfreq
              << "\t nstart = " << 0 << "\t gamma = " << gamma << "\t sexcut = " << sexcut
              << "\t limitone = " << limitone << "\t limittwo = " << limittwo << std::endl;
    const size_t esz = a.mixed ? sizeof(fk_cfloat) : sizeof(fk_cdouble);
    std::cout << "Size of wtilde_array = " << (ncouls * (double)ngpown * esz) / (1024.0 * 1024.0) << " Mbytes"
              << std::endl;
    std::cout << "Size of aqsntemp = " << (ncouls * (double)nb * esz) / (1024.0 * 1024.0) << " Mbytes" << std::endl;
    std::cout << "Size of I_eps_array array = " << (ncouls * (double)ngpown * esz) / (1024.0 * 1024.0) << " Mbytes"
              << std::endl;

    // wx_array[iw] = e_lk - e_n1kq + dw*((iw+1)-2), clamped at to1 (:310-314)
    std::vector<double> wx(nfreq);
    for (int iw = 0; iw < nfreq; ++iw) {
        wx[iw] = e_lk - e_n1kq + dw * ((iw + 1) - 2);
        if (std::abs(wx[iw]) < to1) wx[iw] = to1;
    }
    // column maps; indinv carries the extra entry indinv[ncouls] = ncouls-1
    std::vector<int32_t> inv(ngpown), indinv(ncouls + 1);
    for (int i = 0; i < ngpown; ++i)
        inv[i] = a.random ? (int32_t)((int64_t)i * ncouls / ngpown) : (int32_t)((int64_t)(i + 1) * ncouls / ngpown);
    for (int i = 0; i < ncouls; ++i) indinv[i] = i;
    indinv[ncouls] = ncouls - 1;

    const size_t nbc = (size_t)nb * ncouls, gc = (size_t)ngpown * ncouls;
    fk_options opt = fk_default_options();
    opt.precision = a.mixed ? FK_PREC_MIXED : FK_PREC_FP64;
    opt.mask = a.unmasked ? 0 : 1;
    opt.device = a.device;

    // inputs: host constants (the reference's) or a device fill
    std::vector<fk_cdouble> h64;
    std::vector<fk_cfloat> h32;
    std::vector<double> hv;
    void *d_m = nullptr, *d_n = nullptr, *d_e = nullptr, *d_w = nullptr;
    double *d_v = nullptr;
    const void *pm, *pn, *pe, *pw;
    const double *pv;
    if (!a.random) {
        opt.memory = FK_MEM_HOST;
        hv.assign(ncouls, 1.0);
        // every aqs* entry and every I_eps/wtilde entry is 0.5+0.5i: one buffer serves all four
        const size_t n = std::max(nbc, gc);
        if (a.mixed) {
            h32.assign(n, fk_cfloat{0.5f, 0.5f});
            pm = pn = pe = pw = h32.data();
        } else {
            h64.assign(n, fk_cdouble{0.5, 0.5});
            pm = pn = pe = pw = h64.data();
        }
        pv = hv.data();
    } else {
        opt.memory = FK_MEM_DEVICE;
        hip_check(hipMalloc(&d_m, nbc * esz), "hipMalloc aqsmtemp");
        hip_check(hipMalloc(&d_n, nbc * esz), "hipMalloc aqsntemp");
        hip_check(hipMalloc(&d_e, gc * esz), "hipMalloc I_eps_array");
        hip_check(hipMalloc(&d_w, gc * esz), "hipMalloc wtilde_array");
        hip_check(hipMalloc((void **)&d_v, ncouls * sizeof(double)), "hipMalloc vcoul");
        const int c64 = a.mixed ? 1 : 0;
        const double wlo = a.near_pole ? 3.0 : 0.5, whi = a.near_pole ? 5.0 : 2.0;
        int rc = fk_fill_complex(d_m, (int64_t)nbc, a.seed, 1, -0.5, 0.5, -0.5, 0.5, c64, nullptr);
        if (rc == FK_OK) rc = fk_fill_complex(d_n, (int64_t)nbc, a.seed, 2, -0.5, 0.5, -0.5, 0.5, c64, nullptr);
        if (rc == FK_OK) rc = fk_fill_complex(d_e, (int64_t)gc, a.seed, 3, -0.5, 0.5, -0.5, 0.5, c64, nullptr);
        if (rc == FK_OK) rc = fk_fill_complex(d_w, (int64_t)gc, a.seed, 4, wlo, whi, -0.01, 0.01, c64, nullptr);
        if (rc == FK_OK) rc = fk_fill_real(d_v, ncouls, a.seed, 5, 0.1, 10.0, nullptr);
        if (rc != FK_OK) fail("input fill", rc);
        hip_check(hipDeviceSynchronize(), "input fill");
        pm = d_m; pn = d_n; pe = d_e; pw = d_w; pv = d_v;
    }
    fk_gpp_problem p;
    std::memset(&p, 0, sizeof(p));
    p.number_bands = nb; p.nvband = nv; p.ncouls = ncouls; p.ngpown = ngpown; p.nfreq = nfreq;
    p.wx = wx.data();
    p.aqsmtemp = pm; p.aqsntemp = pn; p.I_eps_array = pe; p.wtilde_array = pw; p.vcoul = pv;
    p.inv_igp_index = inv.data(); p.indinv = indinv.data();
    p.to1 = to1; p.limittwo = limittwo; p.limitone = limitone; p.occ = 1.0;
    p.semantics = a.gppsum == 1 ? FK_GPP_GPPSUM : FK_GPP_CANONICAL;

    if (!a.dump.empty()) dump_instance(a.dump, p, a.mixed, opt.memory == FK_MEM_DEVICE);
    run_and_print(p, opt, a);

    for (void *ptr : {d_m, d_n, d_e, d_w, (void *)d_v})
        if (ptr) (void)hipFree(ptr);
    return 0;
}
