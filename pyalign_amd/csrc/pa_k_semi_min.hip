// Instantiates the DP kernels for locality semi, direction minimize.
#include "pa_kernels.h"

namespace pa {

hipError_t dispatch_semi_min(const LaunchSpec &s, const KArgs &a, hipStream_t st, int *occ) {
    if (s.gap == GAP_LIN)
        return occ ? dispatch_lg<LOC_SEMI, DIR_MIN, GAP_LIN, true>(s, a, st, occ)
                   : dispatch_lg<LOC_SEMI, DIR_MIN, GAP_LIN, false>(s, a, st, occ);
    if (s.gap == GAP_AFF)
        return occ ? dispatch_lg<LOC_SEMI, DIR_MIN, GAP_AFF, true>(s, a, st, occ)
                   : dispatch_lg<LOC_SEMI, DIR_MIN, GAP_AFF, false>(s, a, st, occ);
    return hipErrorInvalidValue;
}

hipError_t dispatch_long_semi_min(int gap, int form, int tr, bool mx, const KArgs &a, const LongArgs &la, int grid, size_t lds,
                                  hipStream_t st) {
    if (gap == GAP_LIN) return dispatch_long_lg<LOC_SEMI, DIR_MIN, GAP_LIN>(form, tr, mx, a, la, grid, lds, st);
    if (gap == GAP_AFF) return dispatch_long_lg<LOC_SEMI, DIR_MIN, GAP_AFF>(form, tr, mx, a, la, grid, lds, st);
    return hipErrorInvalidValue;
}

}  // namespace pa
