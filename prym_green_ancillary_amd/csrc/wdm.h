// wdm.h -- the symmetric WDM text format (cuda/wdmfiles.h:153-222) and its
// optional zstd frame (cuda/zstd_compat.h:56-77,210-333), read by the BCW rank
// tool (third_party/bcw/src/wdm_files.rs:41-148).
//
// Layout: "m n p L b" / Dr / Dc / b lines of V / b lines of curV /
// b(b+1)/2 lines of L sequence values (one packed upper-triangle entry per
// line), space separated canonical residues.  Files whose name ends in .zst
// are written as one zstd stream (level 3).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace pgwio {

bool path_requests_compression(const std::string& path);

// Threads of the WDM writer: PGW_WDM_THREADS, else the host's hardware
// threads, at most 16.
unsigned writer_threads();

// Buffered text sink writing plain bytes or one zstd frame (level 3; with
// threads > 1, libzstd's worker threads compress blocks of that frame).
class TextSink {
public:
    TextSink(const std::string& path, bool compress, unsigned threads = 1);
    ~TextSink();
    void put(const char* data, size_t n);
    void put_u64(uint64_t v);
    void put_char(char c);
    void finish();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// Whole-file reader that undoes a zstd frame when the magic is present.
std::string read_maybe_compressed(const std::string& path);

// Writes `path` via path + ".tmp" and a rename, as save_all_data does
// (cuda/cuprym_cyclic.cu:965-1007).  v and curv are row-major cols x b,
// seq is L x b(b+1)/2 row-major (row t = packed V^T S^t V).  The text is
// formatted on `threads` threads (0: writer_threads()) in pieces that are
// written in order, so every thread count gives the same text (and one
// zstd frame for .zst paths).
void save_wdm_sym(const std::string& path, uint64_t rows, uint64_t cols, uint32_t p,
                  const uint32_t* dr, const uint32_t* dc, const uint32_t* v, const uint32_t* curv,
                  uint32_t b, const uint32_t* seq, uint64_t L, unsigned threads = 0);

// The same file as save_wdm_sym, written while the sequence is computed.
// The constructor starts a writer thread that formats and compresses the
// header, Dr, Dc and the V lines at once (they are known before the loop).
// terms(seq, done) says that rows [0, done) of `seq` (L x b(b+1)/2, the
// caller's array, which must stay alive until finish) hold their final
// values; the writer formats them into the b(b+1)/2 entry lines in memory
// while later rows are still being computed (up to `max_text_bytes` of such
// text; a longer sequence is formatted at the end instead).  finish(curv)
// writes the curV lines and the entry lines, closes the frame and renames
// the file; its bytes equal save_wdm_sym's for the same inputs.  Destroying
// an unfinished stream stops the writer and removes the .tmp file.
class WdmStream {
public:
    WdmStream(const std::string& path, uint64_t rows, uint64_t cols, uint32_t p, const uint32_t* dr,
              const uint32_t* dc, const uint32_t* v, uint32_t b, uint64_t L, unsigned threads = 0,
              uint64_t max_text_bytes = 1ull << 31);
    ~WdmStream();
    WdmStream(const WdmStream&) = delete;
    WdmStream& operator=(const WdmStream&) = delete;
    void terms(const uint32_t* seq, uint64_t done);
    void finish(const uint32_t* curv);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

struct WdmInitial {
    uint64_t rows = 0, cols = 0, length = 0;
    uint32_t prime = 0, block = 0;
    std::vector<uint32_t> dr, dc, v;  // v row-major cols x b
};
// load_wdm_initial_state (cuda/wdmfiles.h:97-149).
WdmInitial load_wdm_initial(const std::string& path);

struct WdmFull : WdmInitial {
    std::vector<uint32_t> curv, seq;  // curv cols x b, seq L x b(b+1)/2
};
WdmFull load_wdm_full(const std::string& path);

}  // namespace pgwio
