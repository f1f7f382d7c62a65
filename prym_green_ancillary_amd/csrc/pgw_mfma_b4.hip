// pgw_mfma_b4.hip -- MFMA sweep instantiations for block width 4
// (one translation unit per width so that they compile in parallel).
#include "pgw_mfma_sweep.h"

namespace pgw {

void launch_dense_mfma_b4(int cfg, uint32_t nl, const DenseSweep& s, const MfmaTable& t,
                           const int8_t* lp, uint32_t* out, const uint32_t* scale, const ModP& m,
                           hipStream_t st) {
    if (nl == 4) dense_mfma_pick<4, 4>(cfg, s, t, lp, out, scale, m, st);
    else dense_mfma_pick<4, 2>(cfg, s, t, lp, out, scale, m, st);
}

uint64_t mfma_group_prepare_b4(int cfg, uint32_t nl, const DenseSweep& s, const MfmaTable& t,
                                const int8_t* lp, uint32_t* out, const uint32_t* scale, const ModP& m,
                                MfmaSweepArgs& a) {
    return nl == 4 ? mfma_group_prepare_t<4, 4>(cfg, s, t, lp, out, scale, m, a)
                   : mfma_group_prepare_t<4, 2>(cfg, s, t, lp, out, scale, m, a);
}

void mfma_group_launch_b4(int cfg, uint32_t nl, bool fwd, const MfmaSweepArgs* jobs,
                           const uint32_t* first, uint32_t njobs, uint64_t blocks, bool persistent, hipStream_t st) {
    if (nl == 4) mfma_group_launch_t<4, 4>(cfg, fwd, jobs, first, njobs, blocks, persistent, st);
    else mfma_group_launch_t<4, 2>(cfg, fwd, jobs, first, njobs, blocks, persistent, st);
}

}  // namespace pgw
