// persist.hip -- persistent-kernel enumeration with device task queues.
#include "qc_internal.h"

namespace qc {

int enumerate_persistent(const Instance* const* inst, const qc_params* params, int count, EnumOut* outs,
                         hipStream_t s) {
    (void)inst;
    (void)params;
    (void)count;
    (void)outs;
    (void)s;
    set_error("persistent mode is not available in this build");
    return QC_EINVAL;
}

}  // namespace qc
