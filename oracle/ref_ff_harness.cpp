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
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <vector>

class GPUComplex;
void achsDtemp_Kernel(int number_bands, int ngpown, int ncouls, int *inv_igp_index, int *indinv,
                      GPUComplex *aqsntemp, GPUComplex *aqsmtemp, GPUComplex *I_epsR_array,
                      double *vcoul, GPUComplex &achsDtemp);

template <class T> static void rd(FILE *f, T *p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
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
