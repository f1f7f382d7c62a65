// pgw_io.cpp -- the WDM writers behind the C ABI (pgw_wdm_write, pgw_wdm_stream_*).
#include "../../include/pgw.h"
#include "pgw_op.h"
#include "wdm.h"

#include <string>
#include <utility>

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

struct pgw_wdm_stream {
    pgwio::WdmStream w;
    template <typename... A>
    explicit pgw_wdm_stream(A&&... a) : w(std::forward<A>(a)...) {}
};

extern "C" int pgw_wdm_stream_open(const char* path, uint64_t rows, uint64_t cols, uint32_t prime,
                                   const uint32_t* h_dr, const uint32_t* h_dc, const uint32_t* h_v,
                                   uint32_t block, uint64_t length, uint32_t threads,
                                   uint64_t max_text_bytes, pgw_wdm_stream** out) {
    return guarded([&] {
        require(path && h_dr && h_dc && h_v && out, "null path, buffer or output pointer");
        require(block >= 1 && block <= 16, "block width must be in 1..16");
        require(prime > 2, "prime must exceed 2");
        require(length >= 1, "Cannot save WDM file with empty sequence list");
        *out = nullptr;
        try {
            *out = new pgw_wdm_stream(std::string(path), rows, cols, prime, h_dr, h_dc, h_v, block, length,
                                      threads, max_text_bytes);
        } catch (const std::bad_alloc&) {
            throw;
        } catch (const std::exception& e) {
            throw Error(PGW_ERR_IO, e.what());
        }
    });
}

extern "C" int pgw_wdm_stream_terms(pgw_wdm_stream* s, const uint32_t* h_seq, uint64_t done) {
    return guarded([&] {
        require(s && h_seq, "null stream or sequence");
        s->w.terms(h_seq, done);
    });
}

extern "C" int pgw_wdm_stream_finish(pgw_wdm_stream* s, const uint32_t* h_curv) {
    return guarded([&] {
        require(s && h_curv, "null stream or curV block");
        try {
            s->w.finish(h_curv);
        } catch (const std::exception& e) {
            throw Error(PGW_ERR_IO, e.what());
        }
    });
}

extern "C" int pgw_wdm_stream_destroy(pgw_wdm_stream* s) {
    delete s;
    return PGW_OK;
}
