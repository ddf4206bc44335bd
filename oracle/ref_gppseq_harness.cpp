// ref_gppseq_harness.cpp -- drives the reference's own compiled gppKerSeq.cpp
// functions on seeded inputs to write golden vectors (TEST INFRASTRUCTURE
// ONLY).
//
// Linked against gppKerSeq.cpp compiled from /root/reference with
// -Dmain=fk_unused_reference_main (see oracle/Makefile), so the functions
// called here are the reference's:
//   flagOCC_solver    gppKerSeq.cpp:76-191   (occupied bands)
//   noflagOCC_solver  gppKerSeq.cpp:193-261  (unoccupied bands, per ig block)
// The loop below is the one of its main (gppKerSeq.cpp:383-472): igmax = igp
// for gppsum = 1 and ncouls otherwise, the occupied bands call
// flagOCC_solver per frequency, the unoccupied ones call noflagOCC_solver
// per 512-wide ig block and frequency, and the sums are scaled and added in
// the same order.  The reference's arrays are [ncouls][nb] and
// [ncouls][ngpown] rows of pointers; the harness builds them from the ABI
// layouts it reads.
//
// Usage: ref_gppseq_harness <input.bin> <output.bin>
//   input : int32 nb, nv, ncouls, ngpown, nw, gppsum; double occ; double wx[nw];
//           complex aqsm[nb*C], aqsn[nb*C], I_eps[G*C], wtilde[G*C];
//           double vcoul[C]; int32 inv_igp_index[G], indinv[C+1]
//   output: complex achtemp[nw], asxtemp[nw], acht_n1_loc[nb]
#include <algorithm>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

typedef std::complex<double> cd;

void flagOCC_solver(double wxt, cd **wtilde_array, int my_igp, int n1, cd **aqsmtemp, cd **aqsntemp,
                    cd **I_eps_array, int gppsum, cd &ssxt, cd &scht, int igmax, int ncouls, int igp);
void noflagOCC_solver(int igbeg, int igend, int igblk, double wxt, cd **wtilde_array, int my_igp, int n1,
                      cd **aqsmtemp, cd **aqsntemp, cd **I_eps_array, int gppsum, cd &ssxt, cd &scht, int igmax,
                      int ncouls, int igp);

template <class T> static void rd(FILE *f, T *p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}
template <class T> static void wr(FILE *f, const T *p, size_t n)
{
    if (n && fwrite(p, sizeof(T), n, f) != n) { fprintf(stderr, "short write\n"); exit(2); }
}

// rows[i] -> the i-th row of a [rows][cols] transposed copy of src[cols][rows]
struct Rows {
    std::vector<cd> data;
    std::vector<cd *> rows;
    Rows(const std::vector<cd> &src, int nsrc, int C)   // src [nsrc][C] -> rows [C][nsrc]
        : data((size_t)C * std::max(nsrc, 1)), rows(C)
    {
        for (int i = 0; i < nsrc; ++i)
            for (int ig = 0; ig < C; ++ig) data[(size_t)ig * nsrc + i] = src[(size_t)i * C + ig];
        for (int ig = 0; ig < C; ++ig) rows[ig] = data.data() + (size_t)ig * std::max(nsrc, 1);
    }
};

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]); return 2; }
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) { perror("input"); return 2; }
    int32_t h[6];
    rd(fi, h, 6);
    const int nb = h[0], nv = h[1], C = h[2], G = h[3], nw = h[4], gppsum = h[5];
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

    Rows m(aqsm, nb, C), n(aqsn, nb, C), e(ieps, G, C), w(wt, G, C);
    std::vector<cd> ach(nw), asx(nw), n1loc(nb);
    for (int n1 = 0; n1 < nb; ++n1) {
        const bool flag_occ = n1 < nv;
        for (int my_igp = 0; my_igp < G; ++my_igp) {
            const int indigp = inv[my_igp];
            int igp = indinv[indigp];
            if (indigp == C) igp = C - 1;
            if (igp > C || igp < 0) continue;
            const int igmax = gppsum == 1 ? igp : C;
            std::vector<cd> ssx_array(nw), sch_array(nw);
            if (flag_occ) {
                for (int iw = 0; iw < nw; ++iw) {
                    cd scht(0, 0), ssxt(0, 0);
                    flagOCC_solver(wx[iw], w.rows.data(), my_igp, n1, m.rows.data(), n.rows.data(), e.rows.data(),
                                   gppsum, ssxt, scht, igmax, C, igp);
                    ssx_array[iw] += ssxt;
                    sch_array[iw] += (double)0.5 * scht;
                }
            } else {
                const int igblk = 512;
                for (int igbeg = 0; igbeg < igmax; igbeg += igblk) {
                    const int igend = std::min(igbeg + igblk - 1, igmax);
                    for (int iw = 0; iw < nw; ++iw) {
                        cd scht(0, 0), ssxt(0, 0);
                        noflagOCC_solver(igbeg, igend, igblk, wx[iw], w.rows.data(), my_igp, n1, m.rows.data(),
                                         n.rows.data(), e.rows.data(), gppsum, ssxt, scht, igmax, C, igp);
                        sch_array[iw] += (double)0.5 * scht;
                    }
                }
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
    fclose(fo);
    return 0;
}
