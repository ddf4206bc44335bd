// fk_ff_split.hip -- the FF contraction recomputed in split BF16 on MFMA (gfx950).
//
// The GEMM of the FF contraction (ffkernel.deslippe.best.f90.back1:402-418),
//     T[n1][m] = sum_ig aqsn[n1][ig] X[m][ig],        X = I_eps as [nfreq*ngpown][ncouls],
// is evaluated from the same FP64 arrays with each real operand rounded to
// FP32 and written as three BF16 slices, x ~ x0 + x1 + x2, where x0 is the
// BF16 rounding of x, x1 that of x - x0, x2 that of x - x0 - x1 (every
// difference is exact in FP32).  The six slice products with i + j <= 2 run
// on v_mfma_f32_32x32x16_bf16 with FP32 accumulation; the three dropped
// products are below 2^-24 of |x||y|.
//
// Accuracy is FP32-class: the operands are rounded to FP32 and T is summed
// in FP32 over one chunk of ig (1024 by default), about 7e-7 normwise
// against FP64 at 256/4096/4096/32.  What runs in FP64 is everything after
// the chunk sums: a workgroup contracts its tile with the coefficients,
//     schDt part[n1] = sum_m cf[n1][m] T[n1][m],   schDv part[n1] = sum_m vg[m] cf[n1][m] T[n1][m],
// and the chunks and column tiles are summed in a fixed order by the
// reduction of fk_ff.hip, so FP32 rounding spans one chunk only.
//
// The complex product as three real GEMMs over K = ncouls (the 3M form):
//     k1 = sum_ig (ar + ai) xr,   k2 = sum_ig ar (xi - xr),   k3 = sum_ig ai (xr + xi),
//     Re T = k1 - k3,   Im T = k1 + k2                              (a = aqsn, x = X)
// with the sums and differences formed in FP64 and rounded to FP32 before
// the split.  Per 32 ig this is 3 x 2 MFMA k-steps of K = 16 where the
// product as two real GEMMs over an interleaved K' = 2 ncouls takes
// 2 x 4; the error stays at the FP32 level of the split.
//
// Tile: 256 bands x 64 columns of one frequency x one ig chunk per
// workgroup of 4 waves; wave w owns 64 bands (2 x 2 blocks of 32x32, one
// accumulator per product: 192 FP32 accumulators).  X is read from HBM
// once: 32 ig per stage, two consecutive ig of a column per lane, split in
// registers into the three B matrices' slices (two ig per dword) and
// written to a double-buffered LDS stage that the four waves share.  Aqsn
// is split once by ff_split_a_kernel into the three A matrices in MFMA
// fragment order, and each wave loads its own fragments from L2 one stage
// ahead.  The MFMAs of a stage run in six groups of 24 (k-step, product);
// the splitting of X and the loads ride between them.
#include "fk_complex.h"
#include "fk_internal.h"

#include <algorithm>

namespace fk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kXcds = 8;
constexpr int kSlices = 3;
constexpr int kProds = 3;                    // k1, k2, k3
constexpr int kStageIg = 32;                 // ig per stage
constexpr int kSteps = 2;                    // MFMA k-steps (K = 16 ig) per stage
constexpr int kTileN = 64;                   // columns per workgroup
constexpr int kTileM = 256;                  // bands per workgroup
constexpr int kWaves = 4;
constexpr int kRowDw = 20;                   // LDS row of one slice: 16 dwords (32 ig) + 4 pad
constexpr int kSliceDw = kTileN * kRowDw;    // one slice of one B matrix for one stage
constexpr int kBufDw = kProds * kSlices * kSliceDw;   // one stage: [product][slice][column][ig pair]
constexpr int kEpiRow = 33;                  // epilogue row: 32 float2 + 1 pad
constexpr int kEpiWave = 64 * kEpiRow;       // float2 per wave
constexpr int kLdsBytes = (2 * kBufDw * 4 > kWaves * kEpiWave * 8) ? 2 * kBufDw * 4 : kWaves * kEpiWave * 8;

// (bf16(lo), bf16(hi)) in one dword, round to nearest even
__device__ __forceinline__ uint32_t cvt_pk_bf16(float lo, float hi)
{
    uint32_t d;
    asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(d) : "v"(lo), "v"(hi));
    return d;
}

// (x, y) -> three slice dwords (x_s in the low half, y_s in the high half),
// s = 0, 1, 2
__device__ __forceinline__ void split3(float x, float y, uint32_t (&d)[kSlices])
{
#pragma unroll
    for (int s = 0; s < kSlices; ++s) {
        d[s] = cvt_pk_bf16(x, y);
        if (s + 1 < kSlices) {
            x -= __uint_as_float(d[s] << 16);
            y -= __uint_as_float(d[s] & 0xffff0000u);
        }
    }
}

__device__ __forceinline__ f32x16 mfma(uint4 a, uint4 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// As[(((q*RF + r)*3 + p)*3 + s)*64 + L] (16 bytes): slice s of A matrix p
// (ar + ai, ar, ai) for fragment row 32r + L%32 and ig = 16q + 8(L/32) ..
// +7, two ig per dword.  Zero past nb and past ncouls.
__global__ void ff_split_a_kernel(const double2 *aqsn, int nb, int C, int nq, int RF, uint4 *As)
{
    const int64_t n = (int64_t)nq * RF * 64;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += stride) {
        const int L = (int)(idx % 64);
        const int r = (int)((idx / 64) % RF);
        const int q = (int)(idx / (64 * RF));
        const int row = 32 * r + (L & 31);
        const int ig0 = 16 * q + 8 * (L >> 5);
        float m[kProds][8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ig = ig0 + e;
            double2 v = make_double2(0.0, 0.0);
            if (row < nb && ig < C) v = aqsn[(size_t)row * C + ig];
            m[0][e] = (float)(v.x + v.y);
            m[1][e] = (float)v.x;
            m[2][e] = (float)v.y;
        }
#pragma unroll
        for (int pr = 0; pr < kProds; ++pr) {
            uint32_t d[4][kSlices];
#pragma unroll
            for (int e = 0; e < 4; ++e) split3(m[pr][2 * e], m[pr][2 * e + 1], d[e]);
#pragma unroll
            for (int sl = 0; sl < kSlices; ++sl)
                As[((((size_t)q * RF + r) * kProds + pr) * kSlices + sl) * 64 + L] =
                    make_uint4(d[0][sl], d[1][sl], d[2][sl], d[3][sl]);
        }
    }
}

struct SplitArgs {
    const uint4 *As;        // split aqsn (ff_split_a_kernel)
    const double2 *X;       // [nf*G][C]
    const double2 *cf;      // [nb][G]
    const double *vg;       // [G]
    double2 *partials;      // [nf][nkc*ngt][nb][2]
    int nb, C, G, nf;
    int RF;                 // fragment rows of As
    int nstages;            // stages over ncouls
    int ns;                 // stages per chunk
    int nkc, ngt, nmt, total;
};

__global__ __launch_bounds__(256, 1) void ff_split_gemm_kernel(SplitArgs a)
{
    __shared__ __align__(16) uint32_t lds[kLdsBytes / 4];

    // workgroup -> tile: each XCD takes a contiguous range of tiles; f runs
    // fastest, so the workgroups that run together on one XCD share one
    // cf block and one chunk of As in its L2
    int t;
    {
        const int b = blockIdx.x, q = a.total / kXcds, r = a.total % kXcds, x = b % kXcds, i = b / kXcds;
        t = (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
    }
    const int f = t % a.nf;
    int rest = t / a.nf;
    const int gt = rest % a.ngt;
    rest /= a.ngt;
    const int mt = rest % a.nmt;
    const int kc = rest / a.nmt;

    const int w = threadIdx.x >> 6, L = threadIdx.x & 63;
    const int s0 = kc * a.ns;
    const int nst = min(a.ns, a.nstages - s0);
    // fragment rows of this wave; a wave past the last band repeats the last
    // two fragments and its results are dropped
    const int r0 = min(8 * mt + 2 * w, a.RF - 2);

    // X elements this thread loads: ig pair ip = L & 15 (ig = stage*32 + 2ip,
    // +1, clamped to the last ig, where the split aqsn is zero) of columns
    // 16w + 4j + (L >> 4) of the tile (j = 0..3, clamped to the last
    // column).  Addresses are 32-bit offsets from a per-wave base.
    const int ip = L & 15;
    const int wcol = min(gt * kTileN + 16 * w, a.G - 1);
    const double2 *xbase = a.X + ((size_t)f * a.G + wcol) * a.C;
    uint32_t xoff[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = min(gt * kTileN + 16 * w + 4 * j + (L >> 4), a.G - 1);
        xoff[j] = (uint32_t)(col - wcol) * (uint32_t)a.C;
    }
    struct XPair { double2 v0, v1; };
    auto load_x = [&](int s, int j) -> XPair {
        const uint32_t ig = (uint32_t)((s0 + s) * kStageIg + 2 * ip);
        const uint32_t last = (uint32_t)(a.C - 1);
        return XPair{xbase[xoff[j] + min(ig, last)], xbase[xoff[j] + min(ig + 1, last)]};
    };
    // split one pair of X elements into the slices of the three B matrices
    // (xr, xi - xr, xr + xi) of LDS stage `buf`
    auto store_x = [&](int buf, int j, const XPair &x) {
        float m[kProds][2];
        m[0][0] = (float)x.v0.x;
        m[0][1] = (float)x.v1.x;
        m[1][0] = (float)(x.v0.y - x.v0.x);
        m[1][1] = (float)(x.v1.y - x.v1.x);
        m[2][0] = (float)(x.v0.x + x.v0.y);
        m[2][1] = (float)(x.v1.x + x.v1.y);
        uint32_t *p = lds + buf * kBufDw + (16 * w + 4 * j + (L >> 4)) * kRowDw + ip;
#pragma unroll
        for (int pr = 0; pr < kProds; ++pr) {
            uint32_t d[kSlices];
            split3(m[pr][0], m[pr][1], d);
#pragma unroll
            for (int sl = 0; sl < kSlices; ++sl) p[(pr * kSlices + sl) * kSliceDw] = d[sl];
        }
    };
    const uint4 *As = a.As + (size_t)r0 * kProds * kSlices * 64 + L;
    const size_t qstride = (size_t)a.RF * kProds * kSlices * 64;   // uint4 per k-step
    auto load_a = [&](int s, int q, int pr, int i, int sl) -> uint4 {
        return As[((size_t)(s0 + s) * kSteps + q) * qstride + ((i * kProds + pr) * kSlices + sl) * 64];
    };

    f32x16 acc[2][2][kProds];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int pr = 0; pr < kProds; ++pr)
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[i][j][pr][e] = 0.f;

    // A fragments of one stage and X of the next stage
    uint4 av[kSteps][kProds][2][kSlices];
    XPair xv[4];
#pragma unroll
    for (int q = 0; q < kSteps; ++q)
#pragma unroll
        for (int pr = 0; pr < kProds; ++pr)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int sl = 0; sl < kSlices; ++sl) av[q][pr][i][sl] = load_a(0, q, pr, i, sl);
#pragma unroll
    for (int j = 0; j < 4; ++j) store_x(0, j, load_x(0, j));
#pragma unroll
    for (int j = 0; j < 4; ++j) xv[j] = load_x(min(1, nst - 1), j);
    __syncthreads();

    for (int s = 0; s < nst; ++s) {
        const uint32_t *buf = lds + (s & 1) * kBufDw;
        const int sn = min(s + 1, nst - 1), sx = min(s + 2, nst - 1);
#pragma unroll
        for (int q = 0; q < kSteps; ++q) {
#pragma unroll
            for (int pr = 0; pr < kProds; ++pr) {
                uint4 b[2][kSlices];   // [column block][slice] of B matrix pr
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int sl = 0; sl < kSlices; ++sl)
                        b[j][sl] = *reinterpret_cast<const uint4 *>(buf + (pr * kSlices + sl) * kSliceDw +
                                                                    (32 * j + (L & 31)) * kRowDw + 8 * q +
                                                                    4 * (L >> 5));
                // slice pairs (A, B): (0,0) (0,1) (0,2) (1,0) (1,1) (2,0)
#pragma unroll
                for (int sa = 0; sa < kSlices; ++sa)
#pragma unroll
                    for (int sb = 0; sb < kSlices - sa; ++sb)
#pragma unroll
                        for (int i = 0; i < 2; ++i)
#pragma unroll
                            for (int j = 0; j < 2; ++j)
                                acc[i][j][pr] = mfma(av[q][pr][i][sa], b[j][sb], acc[i][j][pr]);
                // work between the groups: X of stage s+1 split into the
                // other LDS buffer (last read before the previous barrier),
                // each pair reloaded with stage s+2 at once (groups 1..4),
                // and this group's A fragments reloaded with stage s+1
                const int g = q * kProds + pr;
                if (g >= 1 && g <= 4) {
                    store_x((s + 1) & 1, g - 1, xv[g - 1]);
                    xv[g - 1] = load_x(sx, g - 1);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int sl = 0; sl < kSlices; ++sl) av[q][pr][i][sl] = load_a(sn, q, pr, i, sl);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        __syncthreads();
    }

    // epilogue: T (FP32) -> LDS, one band per thread, FP64 contraction over
    // the tile's columns in ascending order
    float2 *E = reinterpret_cast<float2 *>(lds) + w * kEpiWave;
    const int n1 = mt * kTileM + 64 * w + L;
    const bool own = n1 < a.nb;
    double2 st = make_double2(0.0, 0.0), sv = st;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int row = 32 * i + 8 * (e >> 2) + 4 * (L >> 5) + (e & 3);
                const double k1 = acc[i][j][0][e], k2 = acc[i][j][1][e], k3 = acc[i][j][2][e];
                E[row * kEpiRow + (L & 31)] = make_float2((float)(k1 - k3), (float)(k1 + k2));
            }
        __syncthreads();
        const int c0 = gt * kTileN + 32 * j;
        const int cn = min(32, a.G - c0);
        if (own) {
            const double2 *cfr = a.cf + (size_t)n1 * a.G + c0;
            for (int c = 0; c < cn; ++c) {
                const float2 T = E[L * kEpiRow + c];
                const double2 tt = cmul(cfr[c], make_double2((double)T.x, (double)T.y));
                const double v = a.vg[c0 + c];
                st = cadd(st, tt);
                sv = cadd(sv, make_double2(v * tt.x, v * tt.y));
            }
        }
        __syncthreads();
    }
    if (own) {
        const int mtiles = a.nkc * a.ngt;
        double2 *p = a.partials + (((size_t)f * mtiles + kc * a.ngt + gt) * a.nb + n1) * 2;
        p[0] = st;
        p[1] = sv;
    }
}

// stages per ig chunk for fk_options.config = 0 (default), 1, 2, 3: ig
// chunks of 1024, 256, 512 and 1024.  At 256/4096/4096/32 on MI355X the
// chunk of 1024 runs 4.5 ms with error 6.8e-7 against FP64; 2048 runs 4.33
// ms at 9.6e-7 and 4096 4.30 ms at 1.5e-6 (the FP32 sums grow with the
// chunk), so the default stays at 1024.
constexpr int kSplitNs[] = {32, 8, 16, 32};

}  // namespace

int ff_split_configs() { return (int)(sizeof(kSplitNs) / sizeof(kSplitNs[0])) - 1; }

void ff_split_plan(int nb, int C, int G, int nf, int config, Scratch &ws, FfSplitLayout &L)
{
    L.nstages = div_up(C, kStageIg);
    L.ns = std::min(kSplitNs[std::max(0, std::min(config, ff_split_configs()))], L.nstages);
    L.nkc = div_up(L.nstages, L.ns);
    L.RF = 2 * div_up(nb, 64);
    L.nmt = div_up(nb, kTileM);
    L.ngt = div_up(G, kTileN);
    L.total = L.nkc * L.nmt * L.ngt * nf;
    const size_t nq = (size_t)L.nstages * kSteps;
    L.off_as = ws.plan(nq * L.RF * kProds * kSlices * 64 * sizeof(uint4));
}

int ff_split_launch(const FfSplitLayout &L, Scratch &ws, const double2 *aqsn, const double2 *X, const double2 *cf,
                    const double *vg, double2 *partials, int nb, int C, int G, int nf, KernelTimer &timer,
                    hipStream_t s)
{
    uint4 *As = static_cast<uint4 *>(ws.at(L.off_as));
    const int nq = L.nstages * kSteps;
    {
        const int64_t n = (int64_t)nq * L.RF * 64;
        const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
        hipLaunchKernelGGL(ff_split_a_kernel, dim3(blocks), dim3(256), 0, s, aqsn, nb, C, nq, L.RF, As);
        FK_HIP(hipGetLastError());
    }
    SplitArgs a;
    a.As = As; a.X = X; a.cf = cf; a.vg = vg; a.partials = partials;
    a.nb = nb; a.C = C; a.G = G; a.nf = nf;
    a.RF = L.RF; a.nstages = L.nstages; a.ns = L.ns;
    a.nkc = L.nkc; a.ngt = L.ngt; a.nmt = L.nmt; a.total = L.total;
    FK_TRY(timer.start());
    hipLaunchKernelGGL(ff_split_gemm_kernel, dim3(L.total), dim3(64 * kWaves), 0, s, a);
    FK_HIP(hipGetLastError());
    return timer.stop();
}

}  // namespace fk
