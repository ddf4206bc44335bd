// ref_gpp_harness.cpp -- drives the reference's own compiled GPP functions
// on seeded inputs to write golden vectors (TEST INFRASTRUCTURE ONLY).
//
// Linked against BGW-Kernels/gppKer.cpp compiled from /root/reference with
// -Dmain=fk_unused_reference_main (see oracle/Makefile), so the functions
// called here are the reference's:
//   flagOCC_solver   BGW-Kernels/gppKer.cpp:130-159
//   ssxt_scht_solver BGW-Kernels/gppKer.cpp:19-65 (called by flagOCC_solver)
// The loop below mirrors the reference driver's n1 / my_igp / iw structure,
// index map and scaling (BGW-Kernels/gppKer.cpp:303-439), calling
// flagOCC_solver for every band; its sch sum is the same masked formula as
// the inline unoccupied loop (:394-411), with the product taken as
// (aqsn*conj(aqsm))*(delw*I_eps) instead of ((conj(aqsm)*aqsn)*delw)*I_eps.
//
// Usage: ref_gpp_harness <input.bin> <output.bin>
//   input : int32 nb, nv, ncouls, ngpown, nw; double occ; double wx[nw];
//           complex aqsm[nb*C], aqsn[nb*C], I_eps[G*C], wtilde[G*C];
//           double vcoul[C]; int32 inv_igp_index[G], indinv[C+1]
//   output: complex achtemp[nw], asxtemp[nw], acht_n1_loc[nb],
//           scht[nb][G][nw], ssxt[nb][G][nw] (column sums before scaling;
//           zero for skipped columns)
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <vector>

typedef std::complex<double> cd;

void flagOCC_solver(double wxt, cd *wtilde_array_tmp, int my_igp, int n1, cd *aqsmtemp_arr,
                    cd *aqsntemp_arr, cd *I_eps_array_tmp, cd &ssxt, cd &scht, int igmax,
                    int ncouls, int igp, int number_bands, int ngpown);

template <class T> static void rd(FILE *f, T *p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}
template <class T> static void wr(FILE *f, const T *p, size_t n)
{
    if (n && fwrite(p, sizeof(T), n, f) != n) { fprintf(stderr, "short write\n"); exit(2); }
}

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) { perror("input"); return 2; }
    int32_t h[5];
    rd(fi, h, 5);
    const int nb = h[0], nv = h[1], C = h[2], G = h[3], nw = h[4];
    double occ;
    rd(fi, &occ, 1);
    std::vector<double> wx(nw), vcoul(C);
    std::vector<cd> aqsm((size_t)nb * C), aqsn((size_t)nb * C), ieps((size_t)G * C), wt((size_t)G * C);
    std::vector<int32_t> inv(G), indinv(C + 1);
    rd(fi, wx.data(), nw);
    rd(fi, aqsm.data(), aqsm.size());
    rd(fi, aqsn.data(), aqsn.size());
    rd(fi, ieps.data(), ieps.size());
    rd(fi, wt.data(), wt.size());
    rd(fi, vcoul.data(), C);
    rd(fi, inv.data(), G);
    rd(fi, indinv.data(), C + 1);
    fclose(fi);

    std::vector<cd> ach(nw), asx(nw), n1loc(nb), scht_all((size_t)nb * G * nw), ssxt_all((size_t)nb * G * nw);
    for (int n1 = 0; n1 < nb; ++n1) {
        bool flag_occ = n1 < nv;
        for (int my_igp = 0; my_igp < G; ++my_igp) {
            int indigp = inv[my_igp];
            int igp = indinv[indigp];
            if (indigp == C) igp = C - 1;
            if (igp > C || igp < 0) continue;
            std::vector<cd> ssx_array(nw), sch_array(nw);
            for (int iw = 0; iw < nw; ++iw) {
                cd ssxt(0, 0), scht(0, 0);
                flagOCC_solver(wx[iw], wt.data(), my_igp, n1, aqsm.data(), aqsn.data(), ieps.data(),
                               ssxt, scht, C, C, igp, nb, G);
                scht_all[((size_t)n1 * G + my_igp) * nw + iw] = scht;
                ssxt_all[((size_t)n1 * G + my_igp) * nw + iw] = ssxt;
                ssx_array[iw] += ssxt;
                sch_array[iw] += (double)0.5 * scht;
            }
            if (flag_occ)
                for (int iw = 0; iw < nw; ++iw) asx[iw] += ssx_array[iw] * occ * vcoul[igp];
            for (int iw = 0; iw < nw; ++iw) ach[iw] += sch_array[iw] * vcoul[igp];
            if (nw > 2) n1loc[n1] += sch_array[2] * vcoul[igp];
        }
    }
    FILE *fo = fopen(argv[2], "wb");
    if (!fo) { perror("output"); return 2; }
    wr(fo, ach.data(), nw);
    wr(fo, asx.data(), nw);
    wr(fo, n1loc.data(), nb);
    wr(fo, scht_all.data(), scht_all.size());
    wr(fo, ssxt_all.data(), ssxt_all.size());
    fclose(fo);
    return 0;
}
