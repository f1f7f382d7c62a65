// pgw_mfma_b8.hip -- MFMA sweep instantiations for block width 8
// (one translation unit per width so that they compile in parallel).
#include "pgw_mfma_sweep.h"

namespace pgw {

void launch_dense_mfma_b8(int cfg, uint32_t nl, const DenseSweep& s, const MfmaTable& t,
                           const int8_t* lp, uint32_t* out, const uint32_t* scale, const ModP& m,
                           hipStream_t st) {
    if (nl == 4) dense_mfma_pick<8, 4>(cfg, s, t, lp, out, scale, m, st);
    else dense_mfma_pick<8, 2>(cfg, s, t, lp, out, scale, m, st);
}

}  // namespace pgw
