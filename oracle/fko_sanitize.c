/*
 * fko_sanitize.c -- drives every entry point of the CPU oracle on small,
 * ragged problems so that it can run under AddressSanitizer and
 * UndefinedBehaviorSanitizer, or ThreadSanitizer (TEST INFRASTRUCTURE
 * ONLY: built by `make -C oracle sanitize` and `make -C oracle tsan`, run
 * by tests/test_oracle_sanitize.py).
 *
 * Every array is its own exact-size heap block, so a read or write one
 * element past any input or output is reported.  The cases cover what the
 * parity tests lean on: both column maps (the C++ map reads the extra
 * indinv[ncouls] entry, BGW-Kernels/gppKer.cpp:294-299), an index map that
 * must be rejected, no occupied band, one band, one column, 16 frequencies,
 * partial band ranges, the optional outputs given and NULL, ncouls above
 * 512 for the gppsum = 1 block bounds (gppKerSeq.cpp:438-450), and both FF
 * evaluation variants with and without a supplied schDt matrix.
 *
 * Threads: the extended-precision sums are fixed-order, so each such call
 * is repeated with 1 and 3 OpenMP threads and the outputs must agree bit
 * for bit.  The same driver also builds with ThreadSanitizer (clang and
 * LLVM's OpenMP runtime, whose Archer tool tells TSan about OpenMP's
 * barriers and task joins; gcc's libgomp has no such hooks, so under it
 * TSan would flag every parallel loop).
 *
 * Prints one line per case and "fko_sanitize: all cases passed"; exits 1
 * on a wrong return code or a thread-count mismatch.  Sanitizer findings
 * abort the program (-fno-sanitize-recover=all).
 */
#include "fko.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define TO1 1e-6
#define LIMITTWO 0.25
#define LIMITONE (1.0 / (TO1 * 4.0))

static int failures = 0;

static void check(int ok, const char *what)
{
    if (!ok) {
        fprintf(stderr, "FAIL: %s\n", what);
        ++failures;
    }
}

static void *xalloc(size_t bytes)
{
    void *p = malloc(bytes ? bytes : 1);
    if (!p) {
        fprintf(stderr, "out of memory\n");
        exit(2);
    }
    return p;
}

static fko_z *zarr(int64_t n, uint64_t seed, uint64_t stream, double re_lo, double re_hi, double im_lo,
                   double im_hi)
{
    fko_z *a = (fko_z *)xalloc(sizeof(fko_z) * (size_t)n);
    fko_fill_complex(a, n, seed, stream, re_lo, re_hi, im_lo, im_hi);
    return a;
}

static double *darr(int64_t n, uint64_t seed, uint64_t stream, double lo, double hi)
{
    double *a = (double *)xalloc(sizeof(double) * (size_t)n);
    fko_fill_real(a, n, seed, stream, lo, hi);
    return a;
}

/* column maps: cpp = (i+1) C / G (its last entry is C, which reads the
 * extra indinv[C] = C-1); fortran = i C / G */
static void index_maps(int C, int G, int cpp, int32_t **inv, int32_t **indinv)
{
    *inv = (int32_t *)xalloc(sizeof(int32_t) * (size_t)G);
    *indinv = (int32_t *)xalloc(sizeof(int32_t) * (size_t)(C + 1));
    for (int i = 0; i < G; ++i) (*inv)[i] = (int32_t)((int64_t)(cpp ? i + 1 : i) * C / G);
    for (int i = 0; i < C; ++i) (*indinv)[i] = i;
    (*indinv)[C] = C - 1;
}

typedef struct {
    int nb, nv, C, G, nw, cpp;
    double *wx, *vcoul;
    fko_z *aqsm, *aqsn, *I_eps, *wtilde;
    int32_t *inv, *indinv;
} gpp_case;

static gpp_case gpp_make(int nb, int nv, int C, int G, int nw, int cpp, uint64_t seed, double wt_lo, double wt_hi)
{
    gpp_case g = {nb, nv, C, G, nw, cpp, 0, 0, 0, 0, 0, 0, 0, 0};
    g.wx = (double *)xalloc(sizeof(double) * (size_t)nw);
    for (int iw = 0; iw < nw; ++iw) {         /* e_lk - e_n1kq + dw (iw - 1), clamped at to1 */
        double w = 10.0 - 6.0 + 1.0 * (double)(iw - 1);
        g.wx[iw] = (w < TO1 && w > -TO1) ? TO1 : w;
    }
    g.aqsm = zarr((int64_t)nb * C, seed, 1, -0.5, 0.5, -0.5, 0.5);
    g.aqsn = zarr((int64_t)nb * C, seed, 2, -0.5, 0.5, -0.5, 0.5);
    g.I_eps = zarr((int64_t)G * C, seed, 3, -0.5, 0.5, -0.5, 0.5);
    g.wtilde = zarr((int64_t)G * C, seed, 4, wt_lo, wt_hi, -0.01, 0.01);
    g.vcoul = darr(C, seed, 5, 0.1, 10.0);
    index_maps(C, G, cpp, &g.inv, &g.indinv);
    return g;
}

static void gpp_free(gpp_case *g)
{
    free(g->wx); free(g->vcoul); free(g->aqsm); free(g->aqsn); free(g->I_eps); free(g->wtilde);
    free(g->inv); free(g->indinv);
}

static void run_gpp(const gpp_case *g)
{
    char what[160];
    size_t zw = sizeof(fko_z) * (size_t)g->nw, zb = sizeof(fko_z) * (size_t)(g->nb > 0 ? g->nb : 1);
    fko_z *ach1 = xalloc(zw), *asx1 = xalloc(zw), *ach3 = xalloc(zw), *asx3 = xalloc(zw);
    fko_z *n1a = xalloc(zb), *n1b = xalloc(zb);
    for (int mask = 0; mask <= 1; ++mask) {
        /* extended precision: 1 and 3 threads agree bit for bit */
        int want_n1 = g->nw >= 3;
        int r1 = fko_gpp(g->nb, g->nv, g->C, g->G, g->nw, g->wx, g->aqsm, g->aqsn, g->I_eps, g->wtilde, g->vcoul,
                         g->inv, g->indinv, TO1, LIMITTWO, LIMITONE, 1.0, mask, 1, 1, 0, g->nb, ach1, asx1,
                         want_n1 ? n1a : NULL);
        int r3 = fko_gpp(g->nb, g->nv, g->C, g->G, g->nw, g->wx, g->aqsm, g->aqsn, g->I_eps, g->wtilde, g->vcoul,
                         g->inv, g->indinv, TO1, LIMITTWO, LIMITONE, 1.0, mask, 1, 3, 0, g->nb, ach3, asx3,
                         want_n1 ? n1b : NULL);
        snprintf(what, sizeof what, "fko_gpp extended nb=%d nv=%d C=%d G=%d nw=%d mask=%d", g->nb, g->nv, g->C,
                 g->G, g->nw, mask);
        check(r1 == 0 && r3 == 0, what);
        check(!memcmp(ach1, ach3, zw) && !memcmp(asx1, asx3, zw), what);
        if (want_n1) check(!memcmp(n1a, n1b, sizeof(fko_z) * (size_t)g->nb), what);
        /* reference order, one thread; a partial band range */
        int lo = g->nb / 3, hi = g->nb - g->nb / 4;
        int rr = fko_gpp(g->nb, g->nv, g->C, g->G, g->nw, g->wx, g->aqsm, g->aqsn, g->I_eps, g->wtilde, g->vcoul,
                         g->inv, g->indinv, TO1, LIMITTWO, LIMITONE, 1.0, mask, 0, 1, lo, hi, ach1, asx1,
                         want_n1 ? n1a : NULL);
        check(rr == 0, what);
    }
    printf("fko_gpp nb=%d nv=%d C=%d G=%d nw=%d map=%s: ok\n", g->nb, g->nv, g->C, g->G, g->nw,
           g->cpp ? "cpp" : "fortran");

    /* the static-CH remainder, 1 and 3 threads */
    fko_z s1, s3;
    int a1 = fko_gpp_achstemp(g->nb, g->C, g->G, g->aqsm, g->aqsn, g->I_eps, g->vcoul, g->inv, g->indinv, 1, &s1);
    int a3 = fko_gpp_achstemp(g->nb, g->C, g->G, g->aqsm, g->aqsn, g->I_eps, g->vcoul, g->inv, g->indinv, 3, &s3);
    check(a1 == 0 && a3 == 0 && !memcmp(&s1, &s3, sizeof s1), "fko_gpp_achstemp");

    /* the gppsum = 1 band loop */
    if (g->nw >= 1) {
        int want_n1 = g->nw >= 3;
        int q1 = fko_gpp_gppsum(g->nb, g->nv, g->C, g->G, g->nw, g->wx, g->aqsm, g->aqsn, g->I_eps, g->wtilde,
                                g->vcoul, g->inv, g->indinv, TO1, LIMITTWO, LIMITONE, 1.0, 1, ach1, asx1,
                                want_n1 ? n1a : NULL);
        int q3 = fko_gpp_gppsum(g->nb, g->nv, g->C, g->G, g->nw, g->wx, g->aqsm, g->aqsn, g->I_eps, g->wtilde,
                                g->vcoul, g->inv, g->indinv, TO1, LIMITTWO, LIMITONE, 1.0, 3, ach3, asx3,
                                want_n1 ? n1b : NULL);
        check(q1 == 0 && q3 == 0, "fko_gpp_gppsum");
        check(!memcmp(ach1, ach3, zw) && !memcmp(asx1, asx3, zw), "fko_gpp_gppsum threads");
        if (want_n1) check(!memcmp(n1a, n1b, sizeof(fko_z) * (size_t)g->nb), "fko_gpp_gppsum n1 threads");
    }
    free(ach1); free(asx1); free(ach3); free(asx3); free(n1a); free(n1b);
}

static void run_ff(int nb, int C, int G, int nfreq, uint64_t seed)
{
    fko_z *aqsm = zarr((int64_t)nb * C, seed, 1, -0.5, 0.5, -0.5, 0.5);
    fko_z *aqsn = zarr((int64_t)nb * C, seed, 2, -0.5, 0.5, -0.5, 0.5);
    fko_z *I_eps = zarr((int64_t)nfreq * G * C, seed, 6, -0.5, 0.5, -0.5, 0.5);
    double *vcoul = darr(C, seed, 5, 0.1, 10.0);
    int32_t *inv = (int32_t *)xalloc(sizeof(int32_t) * (size_t)G), *indinv = NULL, *unused = NULL;
    index_maps(C, G, 0, &unused, &indinv);
    free(unused);
    for (int i = 0; i < G; ++i) inv[i] = i;      /* identity map, FFKer_cori/ffKernel.cpp:404-405 */
    size_t zs = sizeof(fko_z) * (size_t)nb * (size_t)nfreq;
    fko_z *t1 = xalloc(zs), *v1 = xalloc(zs), *t3 = xalloc(zs), *v3 = xalloc(zs), a1, a3;
    for (int f = 0; f < nfreq; ++f) {
        int r1 = fko_ff(nb, C, G, nfreq, aqsm, aqsn, I_eps, vcoul, inv, indinv, f, 1, t1, v1, &a1);
        int r3 = fko_ff(nb, C, G, nfreq, aqsm, aqsn, I_eps, vcoul, inv, indinv, f, 3, t3, v3, &a3);
        check(r1 == 0 && r3 == 0, "fko_ff");
        check(!memcmp(t1, t3, zs) && !memcmp(v1, v3, zs) && !memcmp(&a1, &a3, sizeof a1), "fko_ff threads");
    }
    check(fko_ff(nb, C, G, nfreq, aqsm, aqsn, I_eps, vcoul, inv, indinv, 0, 2, t1, NULL, NULL) == 0,
          "fko_ff without the optional outputs");
    printf("fko_ff nb=%d C=%d G=%d nfreq=%d: ok\n", nb, C, G, nfreq);
    free(aqsm); free(aqsn); free(I_eps); free(vcoul); free(inv); free(indinv);
    free(t1); free(v1); free(t3); free(v3);
}

static void run_ff_full(int nb, int nv, int C, int G, int nfreq, int nfe, uint64_t seed)
{
    fko_ff_full_problem p;
    memset(&p, 0, sizeof p);
    p.nb = nb; p.nv = nv; p.ncouls = C; p.ngpown = G; p.nfreq = nfreq; p.nfreqeval = nfe;
    fko_z *aqsm = zarr((int64_t)nb * C, seed, 1, -0.5, 0.5, -0.5, 0.5);
    fko_z *aqsn = zarr((int64_t)nb * C, seed, 2, -0.5, 0.5, -0.5, 0.5);
    fko_z *IR = zarr((int64_t)nfreq * G * C, seed, 6, -0.5, 0.5, -0.5, 0.5);
    fko_z *IA = zarr((int64_t)nfreq * G * C, seed, 7, -0.5, 0.5, -0.5, 0.5);
    double *vcoul = darr(C, seed, 5, 0.1, 10.0);
    int32_t *inv = (int32_t *)xalloc(sizeof(int32_t) * (size_t)G), *indinv = NULL, *unused = NULL;
    index_maps(C, G, 0, &unused, &indinv);
    free(unused);
    for (int i = 0; i < G; ++i) inv[i] = i;
    /* the program's axes (ffkernel.deslippe.best.f90:124-171) */
    double *ekq = (double *)xalloc(sizeof(double) * (size_t)nb);
    double *grid = (double *)xalloc(sizeof(double) * (size_t)nfreq);
    double *pref = (double *)xalloc(sizeof(double) * (size_t)nfreq);
    fko_z *brd = (fko_z *)xalloc(sizeof(fko_z) * (size_t)nfreq);
    for (int n = 0; n < nb; ++n) ekq[n] = -10.0 + n;
    for (int f = 0; f < nfreq; ++f) { grid[f] = 2.0 * f; brd[f].re = 0.0; brd[f].im = 0.1; }
    for (int f = 0; f < nfreq; ++f) pref[f] = f < nfreq - 1 ? (grid[f + 1] - grid[f]) / 3.14 : pref[f - 1];
    pref[0] *= 0.5;
    pref[nfreq - 1] *= 0.5;
    p.aqsm = aqsm; p.aqsn = aqsn; p.I_epsR = IR; p.I_epsA = IA; p.vcoul = vcoul; p.inv_igp_index = inv;
    p.indinv = indinv; p.ekq = ekq; p.dFreqGrid = grid; p.pref = pref; p.dFreqBrd = brd;
    p.pref_zb = 0.5 / 3.14; p.freqevalmin = 0.0; p.freqevalstep = 0.5; p.occ = 1.0; p.achs_freq = 0;

    size_t ze = sizeof(fko_z) * (size_t)nfe, zs = sizeof(fko_z) * (size_t)nb * (size_t)nfreq;
    fko_z *sin = zarr((int64_t)nb * nfreq, seed, 8, -0.5, 0.5, -0.5, 0.5);
    for (int variant = 0; variant <= 1; ++variant)
        for (int given = 0; given <= 1; ++given)
            for (int want_s = 0; want_s <= 1; ++want_s) {
                fko_ff_full_out o;
                memset(&o, 0, sizeof o);
                o.asxDtemp = xalloc(ze); o.achDtemp = xalloc(ze); o.achDtemp_cor = xalloc(ze);
                o.achDtemp_corb = xalloc(ze); o.ach2Dtemp = xalloc(ze);
                o.schDt = want_s ? xalloc(zs) : NULL;
                p.schDt_in = given ? sin : NULL;
                int r = fko_ff_full(&p, variant, 2, &o);
                char what[120];
                snprintf(what, sizeof what, "fko_ff_full variant=%d schDt_in=%d schDt_out=%d", variant, given,
                         want_s);
                check(r == 0, what);
                free(o.asxDtemp); free(o.achDtemp); free(o.achDtemp_cor); free(o.achDtemp_corb);
                free(o.ach2Dtemp); free(o.schDt);
            }
    printf("fko_ff_full nb=%d nv=%d C=%d G=%d nfreq=%d nfreqeval=%d: ok\n", nb, nv, C, G, nfreq, nfe);
    free(aqsm); free(aqsn); free(IR); free(IA); free(vcoul); free(inv); free(indinv);
    free(ekq); free(grid); free(pref); free(brd); free(sin);
}

int main(void)
{
    /* (nb, nv, ncouls, ngpown, nw, cpp map, seed, wtilde re range) */
    const struct { int nb, nv, C, G, nw, cpp; double lo, hi; } gc[] = {
        {7, 3, 37, 6, 3, 1, 0.5, 2.0},     /* ragged, C++ map (reads indinv[C]) */
        {7, 3, 37, 6, 3, 0, 3.0, 5.0},     /* Fortran map, near-pole wtilde */
        {1, 1, 16, 1, 1, 1, 0.5, 2.0},     /* one band, one column, one frequency */
        {9, 0, 50, 11, 5, 0, 0.5, 2.0},    /* no occupied band */
        {5, 5, 33, 5, 16, 1, 3.0, 5.0},    /* every band occupied, 16 frequencies */
        {3, 1, 1030, 4, 3, 1, 0.5, 2.0},   /* ncouls above 512 and 1024: gppsum = 1 block bounds */
        {3, 1, 1024, 2, 3, 1, 3.0, 5.0},   /* columns 512 (the diagonal term) and 1023 */
        {0, 0, 20, 4, 3, 0, 0.5, 2.0},     /* no bands */
    };
    for (size_t k = 0; k < sizeof gc / sizeof gc[0]; ++k) {
        gpp_case g = gpp_make(gc[k].nb, gc[k].nv, gc[k].C, gc[k].G, gc[k].nw, gc[k].cpp, 100 + k, gc[k].lo,
                              gc[k].hi);
        run_gpp(&g);
        gpp_free(&g);
    }

    /* an index map the reference would read out of bounds is rejected */
    {
        gpp_case g = gpp_make(3, 1, 20, 4, 3, 0, 7, 0.5, 2.0);
        g.inv[2] = 25;                       /* > ncouls */
        fko_z a[3], s[3];
        int32_t igp[4];
        check(fko_resolve_igp(g.C, g.G, g.inv, g.indinv, igp) == -1, "fko_resolve_igp out-of-range map");
        check(fko_gpp(g.nb, g.nv, g.C, g.G, g.nw, g.wx, g.aqsm, g.aqsn, g.I_eps, g.wtilde, g.vcoul, g.inv, g.indinv,
                      TO1, LIMITTWO, LIMITONE, 1.0, 1, 1, 2, 0, g.nb, a, s, NULL) == -1,
              "fko_gpp out-of-range map");
        g.inv[2] = -1;
        check(fko_gpp(g.nb, g.nv, g.C, g.G, g.nw, g.wx, g.aqsm, g.aqsn, g.I_eps, g.wtilde, g.vcoul, g.inv, g.indinv,
                      TO1, LIMITTWO, LIMITONE, 1.0, 1, 0, 1, 0, g.nb, a, s, NULL) == -1,
              "fko_gpp negative map entry");
        check(fko_gpp(g.nb, g.nv, g.C, g.G, g.nw, g.wx, g.aqsm, g.aqsn, g.I_eps, g.wtilde, g.vcoul, g.inv, g.indinv,
                      TO1, LIMITTWO, LIMITONE, 1.0, 1, 1, 1, 2, 1, a, s, NULL) == -1,
              "fko_gpp inverted band range");
        gpp_free(&g);
        printf("rejected index maps and band ranges: ok\n");
    }

    /* per-point branches: mask, near-pole, zero */
    {
        fko_z ssx, sch, w = {1.0, 0.0}, e = {0.5, 0.0}, zero = {0.0, 0.0};
        int b1 = fko_gpp_point(0.3, w, e, TO1, LIMITTWO, LIMITONE, &ssx, &sch);
        int b2 = fko_gpp_point(3.0, w, e, TO1, LIMITTWO, LIMITONE, &ssx, &sch);
        int b3 = fko_gpp_point(1.0, zero, zero, TO1, LIMITTWO, LIMITONE, &ssx, &sch);
        check(b1 >= 1 && b1 <= 3 && b2 >= 1 && b2 <= 3 && b3 >= 1 && b3 <= 3, "fko_gpp_point branches");
        for (int ig = 0; ig < 1100; ig += 7)
            for (int igp = 0; igp < 1100; igp += 13) {
                int kind = fko_gppsum_unocc_kind(ig, igp);
                check(kind >= 0 && kind <= 2, "fko_gppsum_unocc_kind range");
            }
        printf("fko_gpp_point, fko_gppsum_unocc_kind: ok\n");
    }

    run_ff(6, 40, 9, 3, 200);
    run_ff(1, 16, 1, 1, 201);
    run_ff(5, 70, 70, 4, 202);
    run_ff_full(6, 2, 24, 6, 5, 7, 300);
    run_ff_full(4, 0, 17, 3, 2, 1, 301);
    run_ff_full(3, 3, 16, 16, 9, 4, 302);

    if (failures) {
        printf("fko_sanitize: %d failures\n", failures);
        return 1;
    }
    printf("fko_sanitize: all cases passed\n");
    return 0;
}
