// fk_ff.hip -- FF frequency contraction (placeholder until the MFMA kernel lands).
#include "fk_internal.h"

using namespace fk;

extern "C" {

size_t ff_kernel_workspace_size(const fk_ff_problem *, const fk_options *) { return 0; }

int ff_kernel(const fk_ff_problem *, const fk_options *, fk_cdouble *, fk_cdouble *, fk_cdouble *)
{
    set_error("ff_kernel: not available in this build");
    return FK_ERR_UNSUPPORTED;
}

}  // extern "C"
