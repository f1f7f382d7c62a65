// pgw_io.cpp -- the WDM writer behind the C ABI (pgw_wdm_write).
#include "../../include/pgw.h"
#include "pgw_op.h"
#include "wdm.h"

using namespace pgw;

extern "C" int pgw_wdm_write(const char* path, uint64_t rows, uint64_t cols, uint32_t prime,
                             const uint32_t* h_dr, const uint32_t* h_dc, const uint32_t* h_v,
                             const uint32_t* h_curv, uint32_t block, const uint32_t* h_seq,
                             uint64_t length, uint32_t threads) {
    return guarded([&] {
        require(path && h_dr && h_dc && h_v && h_curv && h_seq, "null path or buffer");
        require(block >= 1 && block <= 16, "block width must be in 1..16");
        require(prime > 2, "prime must exceed 2");
        try {
            pgwio::save_wdm_sym(path, rows, cols, prime, h_dr, h_dc, h_v, h_curv, block, h_seq, length,
                                threads);
        } catch (const std::exception& e) {
            throw Error(PGW_ERR_IO, e.what());
        }
    });
}
