// fk_internal.h -- shared host-side plumbing of the C ABI: error reporting,
// HIP checks, device scratch and the column-map resolution.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fortkernels.h"

namespace fk {

// Message of the last failing call on this thread.
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

#define FK_HIP(expr)                                                              \
    do {                                                                          \
        hipError_t fk_e_ = (expr);                                                \
        if (fk_e_ != hipSuccess) {                                                \
            ::fk::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(fk_e_), \
                            __FILE__, __LINE__);                                  \
            return FK_ERR_HIP;                                                    \
        }                                                                         \
    } while (0)

#define FK_TRY(expr)                    \
    do {                                \
        int fk_rc_ = (expr);            \
        if (fk_rc_ != FK_OK) return fk_rc_; \
    } while (0)

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
inline int div_up(int x, int a) { return (x + a - 1) / a; }

// Column map of BGW-Kernels/gppKer.cpp:322-328: igp = indinv[inv_igp_index[my_igp]],
// indigp == ncouls -> ncouls-1, igp > ncouls or igp < 0 -> -1 (column skipped).
// Returns FK_ERR_INDEX where the reference would read out of bounds
// (indigp outside [0, ncouls], or igp == ncouls).
int resolve_columns(int ncouls, int ngpown, const int32_t *inv_igp_index, const int32_t *indinv,
                    std::vector<int32_t> &igp_out);

// Device scratch: a caller-provided workspace or a stream-ordered allocation
// released by the destructor (stream-ordered as well).
class Scratch {
public:
    Scratch() = default;
    ~Scratch();
    Scratch(const Scratch &) = delete;
    Scratch &operator=(const Scratch &) = delete;
    // Carve `bytes` (256-aligned offsets).  Call plan() for every piece, then
    // acquire(), then ptr() in the same order.
    size_t plan(size_t bytes) { size_t off = total_; total_ = align_up(total_ + bytes, 256); return off; }
    size_t total() const { return total_; }
    int acquire(void *user, size_t user_bytes, hipStream_t stream);
    void *at(size_t offset) const { return static_cast<char *>(base_) + offset; }

private:
    size_t total_ = 0;
    void *base_ = nullptr;
    bool owned_ = false;
    hipStream_t stream_ = nullptr;
};

// Timing of one kernel with HIP events on its stream (only when requested).
class KernelTimer {
public:
    KernelTimer(float *out, hipStream_t s);
    ~KernelTimer();
    int start();
    int stop();
    int finish();   // synchronise and write the elapsed ms

private:
    float *out_;
    hipStream_t s_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

// FP64 row contraction on the device, shared by ff_kernel and the per-band
// output of gpp_kernel:
//   out[n1] = vscale * sum_m vcoul[igp_m] conj(aqsm[n1][igp_m]) sum_ig aqsn[n1][ig] X[m][ig]
// for one X[G][C] (igp_m = cols[m], -1 = skipped column).  Asynchronous on
// `s`; its scratch is `ws` (at least ff_rows_weighted_bytes) or, with
// ws = NULL, a stream-ordered allocation.
size_t ff_rows_weighted_bytes(int nb, int C, int G);
int ff_rows_weighted(int nb, int C, int G, const double2 *aqsm, const double2 *aqsn, const double2 *X,
                     const double *vcoul, const std::vector<int32_t> &cols, double vscale, double2 *out,
                     void *ws, size_t ws_bytes, hipStream_t s);

// Fixed-order sum of GPP per-tile partials [tiles][nw][2] into achtemp and
// asxtemp (asxtemp scaled by asx_scale).
// out = the fixed-order sum of n partials (one workgroup).
int gpp_sum_launch(const double2 *partials, int n, double2 *out, hipStream_t s);

int gpp_reduce_launch(const double2 *partials, int tiles, int nw, double asx_scale, double2 *achtemp,
                      double2 *asxtemp, hipStream_t s);

// The gppsum = 1 form of the GPP band loop (fk_gpp_sym.hip): the occupied
// loop's stop per column (when nvband > 0) into d_stop, then the tile kernel
// over c[Kpad][ldc] (the canonical coefficient matrix) writing per-tile
// partials [tiles_m * tiles_n][nfreq][2] for gpp_reduce_launch.
// kernel_ms (optional) times the tile kernel.
int gpp_sym_launch(const fk_gpp_problem *p, const double2 *c, int ldc, int Kpad, int nvpad, const int *d_igp,
                   int *d_stop, const double2 *aqsm, const double2 *aqsn, const double2 *ieps, const double2 *wt,
                   double2 *partials, int tiles_m, int tiles_n, float *kernel_ms, hipStream_t s);

// Mixed-precision GPP on int8 MFMA (complex64 inputs): workspace layout,
// size limits, and the launch of its three kernels plus the reduction.
struct OzLayout {
    int ldm, ldn, Kb, tiles_m, tiles_n;   // Kb: bands padded to the GEMM step
    size_t off_as, off_bs, off_ea, off_eb, off_cocc, off_bmat, off_part, off_achs;
};
void gpp_oz_plan(const fk_gpp_problem *p, Scratch &ws, OzLayout &L);
int gpp_oz_check(const fk_gpp_problem *p);
int gpp_oz_launch(const fk_gpp_problem *p, const fk_options *o, const OzLayout &L, Scratch &ws, const int *d_igp,
                  const float2 *aqsm, const float2 *aqsn, const float2 *ieps, const float2 *wt, const double *vcoul,
                  double2 *d_ach, double2 *d_asx, double2 *d_achs, hipStream_t s);

// Batches of the int8-split path: each problem's kernel arguments as an
// opaque record of gpp_oz_batch_record_bytes() bytes (written to `record`,
// host memory) and its block counts; gpp_oz_batch_launch runs the prep,
// GEMM and pole kernels of n records (device memory) as one launch each,
// first_* holding each problem's first block (n + 1 ints, device memory).
// The pole partials of a problem are [blocks.pole][nfreq][2] at the
// layout's off_part, for the batched reduction.
struct OzBlocks { int prep, gemm, pole; };
size_t gpp_oz_batch_record_bytes();
OzBlocks gpp_oz_batch_record(const fk_gpp_problem *p, const fk_options *o, const OzLayout &L, const Scratch &ws,
                             const int *d_igp, void *record);
int gpp_oz_batch_launch(const void *records, const int *first_prep, const int *first_gemm, const int *first_pole,
                        int n, int total_prep, int total_gemm, int total_pole, hipStream_t s);

// FF contraction in split BF16 on MFMA (FK_PREC_MIXED): workspace layout and
// the launch of the slice pre-pass and the GEMM.  The GEMM writes FP64
// partials [nf][nkc*ngt][nb][2] (schDt part, schDv part) for the reduction
// of fk_ff.hip; `timer` brackets the GEMM kernel.
struct FfSplitLayout {
    int nstages, ns, nkc, RF, nmt, ngt, total;
    size_t off_as;
};
int ff_split_configs();
void ff_split_plan(int nb, int C, int G, int nf, int config, Scratch &ws, FfSplitLayout &L);
int ff_split_launch(const FfSplitLayout &L, Scratch &ws, const double2 *aqsn, const double2 *X, const double2 *cf,
                    const double *vg, double2 *partials, int nb, int C, int G, int nf, KernelTimer &timer,
                    hipStream_t s);

// The FF contraction of one I_eps array (fk_ff.hip): ff_plan carves its
// workspace out of `ws` (several plans can share one Scratch), ff_launch
// runs the coefficient kernel, the GEMM (FP64 MFMA, or split BF16 for
// FK_PREC_MIXED) and the fixed-order reduction into schDt / schDv (either
// may be NULL) and, if achs is non-NULL, achsDtemp at p->achs_freq.
// options->kernel_ms times the GEMM kernel.  upload_cols = false: the column
// map is already in the layout's igp slot (a caller that captures the
// launches into a graph uploads it first).
// fk_options.config of the FP64 FF GEMM with 16-row tiles (for few bands)
constexpr int kFfConfigFewBands = 3;
struct FfLayout {
    int cfg, tiles_m, mtiles, nks;
    int ksplit = 1;         // ig parts per column block (few tiles, or the few-band kernel)
    int kper = 0;           // k-steps per part (the FP64 fused tiles with ksplit > 1)
    bool split;             // FK_PREC_MIXED: split BF16 (fk_ff_split.hip)
    FfSplitLayout sp;
    size_t off_cf, off_vg, off_igp, off_part, off_achs;
};
int ff_fp64_configs();   // FP64 tile configurations (fk_options.config 1..this)
FfLayout ff_plan(const fk_ff_problem *p, const fk_options *o, Scratch &ws);
int ff_check(const fk_ff_problem *p, const fk_options *o);
int ff_launch(const fk_ff_problem *p, const fk_options *o, const FfLayout &L, Scratch &ws,
              const std::vector<int32_t> &cols, const double2 *aqsm, const double2 *aqsn, const double2 *X,
              const double *vcoul, double2 *schDt, double2 *schDv, double2 *achs, double vscale, hipStream_t s,
              bool upload_cols = true);

}  // namespace fk
