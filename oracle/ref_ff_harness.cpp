// ref_ff_harness.cpp -- drives the reference's compiled FF static-CH kernel
// on seeded inputs to write golden vectors (TEST INFRASTRUCTURE ONLY).
//
// Linked against FFKer_cori/ffKernel.cpp compiled from /root/reference with
// -Dmain=fk_unused_reference_main (see oracle/Makefile); the function called
// is the reference's achsDtemp_Kernel (FFKer_cori/ffKernel.cpp:107-131),
// which reads frequency index 1 of I_epsR_array (:122).  GPUComplex is the
// reference's {double re, im} class (FFKer_cori/Complex.h); it is only
// passed by pointer here, so the harness names it without its definition.
//
// Usage: ref_ff_harness <input.bin> <output.bin>
//   input : int32 nb, ncouls, ngpown, nfreq; complex aqsm[nb*C], aqsn[nb*C],
//           I_eps[nfreq*G*C]; double vcoul[C]; int32 inv_igp_index[G],
//           indinv[C]
//   output: complex achsDtemp
//
// Usage: ref_ff_harness full <input.bin> <output.bin>
//   runs the four kernel functions in the order of the program's main
//   (:485-500), one OpenMP thread, with main's aliasing (achDtemp_Kernel
//   writes into asxDtemp, :495) and zero-initialised work arrays (:455-466).
//   input : int32 nb, nv, ncouls, ngpown, nfreq, nfreqeval;
//           double freqevalmin, freqevalstep, occ, pref_zb;
//           complex aqsm[nb*C], aqsn[nb*C], I_epsR[nfreq*G*C],
//           I_epsA[nfreq*G*C]; double vcoul[C], ekq[nb], dFreqGrid[nfreq],
//           pref[nfreq]; complex dFreqBrd[nfreq], schDt_matrix[nb*nfreq];
//           int32 inv_igp_index[G], indinv[C]
//   output: complex achsDtemp, asxDtemp[nfreqeval], achDtemp[nfreqeval],
//           ach2Dtemp[nfreqeval], achDtemp_cor[nfreqeval],
//           achDtemp_corb[nfreqeval]
#include <omp.h>

#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <string>
#include <vector>

class GPUComplex;
void achsDtemp_Kernel(int number_bands, int ngpown, int ncouls, int *inv_igp_index, int *indinv,
                      GPUComplex *aqsntemp, GPUComplex *aqsmtemp, GPUComplex *I_epsR_array,
                      double *vcoul, GPUComplex &achsDtemp);
void asxDtemp_Kernel(int nvband, int nfreqeval, int ncouls, int ngpown, int numThreads, int nFreq, double freqevalmin,
                     double freqevalstep, double occ, double *ekq, double *dFreqGrid, int *inv_igp_index, int *indinv,
                     GPUComplex *aqsmtemp, GPUComplex *aqsntemp, double *vcoul, GPUComplex *I_epsR_array,
                     GPUComplex *ssxDittt, GPUComplex *asxDtemp);
void achDtemp_Kernel(int number_bands, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                     double freqevalmin, double freqevalstep, double *ekq, double pref_zb, double *pref,
                     double *dFreqGrid, GPUComplex *dFreqBrd, GPUComplex *schDt_matrix, GPUComplex *schDi,
                     GPUComplex *schDi_cor, GPUComplex *sch2Di, GPUComplex *achDtemp);
void achDtemp_cor_Kernel(int number_bands, int nvband, int nfreqeval, int ncouls, int ngpown, int nFreq,
                         double freqevalmin, double freqevalstep, double *ekq, double *dFreqGrid, int *inv_igp_index,
                         int *indinv, GPUComplex *aqsmtemp, GPUComplex *aqsntemp, double *vcoul,
                         GPUComplex *I_epsR_array, GPUComplex *I_epsA_array, GPUComplex *schDi_cor,
                         GPUComplex *schDi_corb, GPUComplex *sch2Di, GPUComplex *ach2Dtemp, GPUComplex *achDtemp_cor,
                         GPUComplex *achDtemp_corb);

template <class T> static void rd(FILE *f, T *p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}

#define GC(v) reinterpret_cast<GPUComplex *>((v).data())

static int run_full(const char *in, const char *out)
{
    FILE *fi = fopen(in, "rb");
    if (!fi) { perror("input"); return 2; }
    int32_t h[6];
    double c[4];
    rd(fi, h, 6);
    rd(fi, c, 4);
    const int nb = h[0], nv = h[1], C = h[2], G = h[3], nf = h[4], ne = h[5];
    std::vector<double> aqsm((size_t)2 * nb * C), aqsn((size_t)2 * nb * C), iR((size_t)2 * nf * G * C),
        iA((size_t)2 * nf * G * C), vcoul(C), ekq(nb), grid(nf), pref(nf), brd((size_t)2 * nf),
        schdt((size_t)2 * nb * nf);
    std::vector<int32_t> inv(G), indinv(C);
    rd(fi, aqsm.data(), aqsm.size());
    rd(fi, aqsn.data(), aqsn.size());
    rd(fi, iR.data(), iR.size());
    rd(fi, iA.data(), iA.size());
    rd(fi, vcoul.data(), vcoul.size());
    rd(fi, ekq.data(), ekq.size());
    rd(fi, grid.data(), grid.size());
    rd(fi, pref.data(), pref.size());
    rd(fi, brd.data(), brd.size());
    rd(fi, schdt.data(), schdt.size());
    rd(fi, inv.data(), G);
    rd(fi, indinv.data(), C);
    fclose(fi);
    omp_set_num_threads(1);
    std::vector<double> achs(2, 0.0), ssx(2, 0.0), schDi(2 * ne, 0.0), schDi_cor(2 * ne, 0.0),
        schDi_corb(2 * ne, 0.0), sch2Di(2 * ne, 0.0), asx(2 * ne, 0.0), ach(2 * ne, 0.0), ach2(2 * ne, 0.0),
        cor(2 * ne, 0.0), corb(2 * ne, 0.0);
    achsDtemp_Kernel(nb, G, C, inv.data(), indinv.data(), GC(aqsn), GC(aqsm), GC(iR), vcoul.data(), *GC(achs));
    asxDtemp_Kernel(nv, ne, C, G, 1, nf, c[0], c[1], c[2], ekq.data(), grid.data(), inv.data(), indinv.data(),
                    GC(aqsm), GC(aqsn), vcoul.data(), GC(iR), GC(ssx), GC(asx));
    achDtemp_Kernel(nb, nv, ne, C, G, nf, c[0], c[1], ekq.data(), c[3], pref.data(), grid.data(), GC(brd), GC(schdt),
                    GC(schDi), GC(schDi_cor), GC(sch2Di), GC(asx));
    achDtemp_cor_Kernel(nb, nv, ne, C, G, nf, c[0], c[1], ekq.data(), grid.data(), inv.data(), indinv.data(),
                        GC(aqsm), GC(aqsn), vcoul.data(), GC(iR), GC(iA), GC(schDi_cor), GC(schDi_corb), GC(sch2Di),
                        GC(ach2), GC(cor), GC(corb));
    FILE *fo = fopen(out, "wb");
    if (!fo) { perror("output"); return 2; }
    fwrite(achs.data(), sizeof(double), 2, fo);
    for (const std::vector<double> *v : {&asx, &ach, &ach2, &cor, &corb}) fwrite(v->data(), sizeof(double), v->size(), fo);
    fclose(fo);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 4 && std::string(argv[1]) == "full") return run_full(argv[2], argv[3]);
    if (argc != 3) { fprintf(stderr, "usage: %s [full] in.bin out.bin\n", argv[0]); return 2; }
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) { perror("input"); return 2; }
    int32_t h[4];
    rd(fi, h, 4);
    const int nb = h[0], C = h[1], G = h[2], nf = h[3];
    std::vector<double> aqsm((size_t)2 * nb * C), aqsn((size_t)2 * nb * C), ieps((size_t)2 * nf * G * C), vcoul(C);
    std::vector<int32_t> inv(G), indinv(C);
    rd(fi, aqsm.data(), aqsm.size());
    rd(fi, aqsn.data(), aqsn.size());
    rd(fi, ieps.data(), ieps.size());
    rd(fi, vcoul.data(), C);
    rd(fi, inv.data(), G);
    rd(fi, indinv.data(), C);
    fclose(fi);
    double out[2] = {0, 0};
    achsDtemp_Kernel(nb, G, C, inv.data(), indinv.data(), reinterpret_cast<GPUComplex *>(aqsn.data()),
                     reinterpret_cast<GPUComplex *>(aqsm.data()), reinterpret_cast<GPUComplex *>(ieps.data()),
                     vcoul.data(), *reinterpret_cast<GPUComplex *>(out));
    FILE *fo = fopen(argv[2], "wb");
    if (!fo) { perror("output"); return 2; }
    fwrite(out, sizeof(double), 2, fo);
    fclose(fo);
    return 0;
}
