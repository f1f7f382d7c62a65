// wdm.cpp -- WDM text writer/reader with an optional zstd frame.

#include "wdm.h"

#include <algorithm>
#include <charconv>
#include <condition_variable>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

// libzstd's stable streaming API (zstd.h is not part of this image, the
// shared library libzstd.so.1 is): opaque stream types and buffer structs.
extern "C" {
struct ZSTD_CCtx_s;
struct ZSTD_DCtx_s;
typedef struct ZSTD_CCtx_s ZSTD_CStream;
typedef struct ZSTD_DCtx_s ZSTD_DStream;
typedef struct {
    const void* src;
    size_t size;
    size_t pos;
} ZSTD_inBuffer;
typedef struct {
    void* dst;
    size_t size;
    size_t pos;
} ZSTD_outBuffer;
ZSTD_CStream* ZSTD_createCStream(void);
size_t ZSTD_freeCStream(ZSTD_CStream*);
size_t ZSTD_initCStream(ZSTD_CStream*, int level);
size_t ZSTD_compressStream(ZSTD_CStream*, ZSTD_outBuffer*, ZSTD_inBuffer*);
size_t ZSTD_endStream(ZSTD_CStream*, ZSTD_outBuffer*);
size_t ZSTD_CStreamOutSize(void);
// advanced one-frame streaming with worker threads (stable since zstd 1.4.0):
// parameter ids and end directives as libzstd defines them
ZSTD_CStream* ZSTD_createCCtx(void);
size_t ZSTD_freeCCtx(ZSTD_CStream*);
size_t ZSTD_CCtx_setParameter(ZSTD_CStream*, int param, int value);
size_t ZSTD_compressStream2(ZSTD_CStream*, ZSTD_outBuffer*, ZSTD_inBuffer*, int end_op);
ZSTD_DStream* ZSTD_createDStream(void);
size_t ZSTD_freeDStream(ZSTD_DStream*);
size_t ZSTD_initDStream(ZSTD_DStream*);
size_t ZSTD_decompressStream(ZSTD_DStream*, ZSTD_outBuffer*, ZSTD_inBuffer*);
size_t ZSTD_DStreamOutSize(void);
unsigned ZSTD_isError(size_t);
const char* ZSTD_getErrorName(size_t);
}

namespace pgwio {

namespace {
constexpr int kZstdCompressionLevel = 100;  // ZSTD_c_compressionLevel
constexpr int kZstdNbWorkers = 400;         // ZSTD_c_nbWorkers
constexpr int kZstdContinue = 0, kZstdEnd = 2;
}  // namespace

unsigned writer_threads() {
    if (const char* env = std::getenv("PGW_WDM_THREADS")) {
        const long t = std::strtol(env, nullptr, 10);
        if (t >= 1 && t <= 256) return (unsigned)t;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return hw == 0 ? 1u : (hw > 16 ? 16u : hw);
}

static bool ends_with(const std::string& s, const std::string& suf) {
    return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

bool path_requests_compression(const std::string& path) {
    return ends_with(path, ".zst") || ends_with(path, ".zst.tmp");
}

struct TextSink::Impl {
    FILE* f = nullptr;
    bool compress = false;
    ZSTD_CStream* zs = nullptr;  // a ZSTD_CCtx: one frame, level 3, worker threads when libzstd has them
    std::vector<char> buf, zout;
    size_t used = 0;
    std::string path;

    void write_raw(const void* data, size_t n) {
        if (n && std::fwrite(data, 1, n, f) != n)
            throw std::runtime_error("failed while writing " + path);
    }
    void flush_buf() {
        if (!used) return;
        if (!compress) {
            write_raw(buf.data(), used);
        } else {
            ZSTD_inBuffer in{buf.data(), used, 0};
            while (in.pos < in.size) {
                ZSTD_outBuffer out{zout.data(), zout.size(), 0};
                size_t rc = ZSTD_compressStream2(zs, &out, &in, kZstdContinue);
                if (ZSTD_isError(rc)) throw std::runtime_error(std::string("zstd: ") + ZSTD_getErrorName(rc));
                write_raw(zout.data(), out.pos);
            }
        }
        used = 0;
    }
};

TextSink::TextSink(const std::string& path, bool compress, unsigned threads) : impl_(new Impl) {
    impl_->path = path;
    impl_->compress = compress;
    impl_->f = std::fopen(path.c_str(), "wb");
    if (!impl_->f) throw std::runtime_error("failed to open " + path);
    impl_->buf.resize(1 << 22);
    if (compress) {
        impl_->zs = ZSTD_createCCtx();
        if (!impl_->zs) throw std::runtime_error("failed to allocate a zstd stream");
        size_t rc = ZSTD_CCtx_setParameter(impl_->zs, kZstdCompressionLevel, 3);
        if (ZSTD_isError(rc)) throw std::runtime_error(std::string("zstd init: ") + ZSTD_getErrorName(rc));
        // worker threads compress blocks of the same frame; a libzstd built
        // without them refuses the parameter and compresses on this thread
        if (threads > 1) (void)ZSTD_CCtx_setParameter(impl_->zs, kZstdNbWorkers, (int)threads);
        impl_->zout.resize(ZSTD_CStreamOutSize());
    }
}

TextSink::~TextSink() {
    if (impl_->zs) ZSTD_freeCCtx(impl_->zs);
    if (impl_->f) std::fclose(impl_->f);
}

void TextSink::put(const char* data, size_t n) {
    while (n) {
        size_t room = impl_->buf.size() - impl_->used;
        size_t take = n < room ? n : room;
        std::memcpy(impl_->buf.data() + impl_->used, data, take);
        impl_->used += take;
        data += take;
        n -= take;
        if (impl_->used == impl_->buf.size()) impl_->flush_buf();
    }
}

void TextSink::put_u64(uint64_t v) {
    char tmp[24];
    auto r = std::to_chars(tmp, tmp + sizeof(tmp), v);
    put(tmp, (size_t)(r.ptr - tmp));
}

void TextSink::put_char(char c) { put(&c, 1); }

void TextSink::finish() {
    impl_->flush_buf();
    if (impl_->compress) {
        for (;;) {
            ZSTD_outBuffer out{impl_->zout.data(), impl_->zout.size(), 0};
            ZSTD_inBuffer none{nullptr, 0, 0};
            size_t left = ZSTD_compressStream2(impl_->zs, &out, &none, kZstdEnd);
            if (ZSTD_isError(left)) throw std::runtime_error(std::string("zstd end: ") + ZSTD_getErrorName(left));
            impl_->write_raw(impl_->zout.data(), out.pos);
            if (left == 0) break;
        }
    }
    if (std::fclose(impl_->f) != 0) {
        impl_->f = nullptr;
        throw std::runtime_error("failed to close " + impl_->path);
    }
    impl_->f = nullptr;
}

std::string read_maybe_compressed(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("failed to open " + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    std::string raw = ss.str();
    static const unsigned char magic[4] = {0x28, 0xb5, 0x2f, 0xfd};
    if (raw.size() < 4 || std::memcmp(raw.data(), magic, 4) != 0) return raw;
    ZSTD_DStream* ds = ZSTD_createDStream();
    if (!ds) throw std::runtime_error("failed to allocate a zstd decompression stream");
    ZSTD_initDStream(ds);
    std::string text;
    std::vector<char> out(ZSTD_DStreamOutSize());
    ZSTD_inBuffer zin{raw.data(), raw.size(), 0};
    while (zin.pos < zin.size) {
        ZSTD_outBuffer zo{out.data(), out.size(), 0};
        size_t rc = ZSTD_decompressStream(ds, &zo, &zin);
        if (ZSTD_isError(rc)) {
            ZSTD_freeDStream(ds);
            throw std::runtime_error(std::string("zstd: ") + ZSTD_getErrorName(rc));
        }
        text.append(out.data(), zo.pos);
    }
    for (;;) {  // drain buffered output
        ZSTD_outBuffer zo{out.data(), out.size(), 0};
        size_t rc = ZSTD_decompressStream(ds, &zo, &zin);
        if (ZSTD_isError(rc)) break;
        text.append(out.data(), zo.pos);
        if (zo.pos == 0) break;
    }
    ZSTD_freeDStream(ds);
    return text;
}

// values i in [begin, end) of a line, each preceded by a space except the line's first
static void format_range(const uint32_t* v, uint64_t begin, uint64_t end, uint64_t stride, uint32_t p,
                         std::string& out) {
    out.clear();
    out.reserve((end - begin) * 11);
    char tmp[16];
    for (uint64_t i = begin; i < end; ++i) {
        if (i) out.push_back(' ');
        auto r = std::to_chars(tmp, tmp + sizeof(tmp), p ? v[i * stride] % p : v[i * stride]);
        out.append(tmp, (size_t)(r.ptr - tmp));
    }
}

// One text line of `count` values (p = 0: written as stored).  Long lines are
// formatted in pieces, `threads` pieces (a window) at a time on worker
// threads, while this thread hands the previous window to the sink (zstd
// compresses on the calling thread): pieces go out in order, so the bytes are
// those of a serial writer.
static void put_line(TextSink& s, const uint32_t* v, uint64_t count, uint64_t stride, uint32_t p,
                     unsigned threads) {
    constexpr uint64_t kPiece = 1ull << 18;  // values per piece (about 2.8 MB of text)
    const uint64_t window = kPiece * threads;
    if (threads <= 1 || count <= kPiece) {
        std::string piece;
        for (uint64_t b = 0; b < count; b += kPiece) {
            format_range(v, b, std::min(count, b + kPiece), stride, p, piece);
            s.put(piece.data(), piece.size());
        }
        s.put_char('\n');
        return;
    }
    std::vector<std::string> bufs[2] = {std::vector<std::string>(threads), std::vector<std::string>(threads)};
    std::vector<std::thread> pool;
    auto launch = [&](uint64_t w0, std::vector<std::string>& dst) {  // format window w0 in the background
        for (unsigned q = 0; q < threads; ++q) {
            const uint64_t b = w0 + q * kPiece;
            if (b >= count) break;
            pool.emplace_back([&, b, q] { format_range(v, b, std::min(count, b + kPiece), stride, p, dst[q]); });
        }
    };
    auto join = [&] {
        for (auto& t : pool) t.join();
        pool.clear();
    };
    launch(0, bufs[0]);
    join();
    for (uint64_t w0 = 0, k = 0; w0 < count; w0 += window, k ^= 1) {
        if (w0 + window < count) launch(w0 + window, bufs[k ^ 1]);
        try {
            for (unsigned q = 0; q < threads && w0 + q * kPiece < count; ++q)
                s.put(bufs[k][q].data(), bufs[k][q].size());
        } catch (...) {
            join();
            throw;
        }
        join();
    }
    s.put_char('\n');
}

// "m n p L b", Dr, Dc and the b lines of V: everything known before the loop
static void put_prefix(TextSink& s, uint64_t rows, uint64_t cols, uint32_t p, const uint32_t* dr,
                       const uint32_t* dc, const uint32_t* v, uint32_t b, uint64_t L, unsigned threads) {
    s.put_u64(rows); s.put_char(' ');
    s.put_u64(cols); s.put_char(' ');
    s.put_u64(p); s.put_char(' ');
    s.put_u64(L); s.put_char(' ');
    s.put_u64(b); s.put_char('\n');
    // preconditioners are written as stored (cuda/wdmfiles.h:195-199)
    put_line(s, dr, rows, 1, 0, threads);
    put_line(s, dc, cols, 1, 0, threads);
    for (uint32_t c = 0; c < b; ++c) put_line(s, v + c, cols, b, p, threads);
}

void save_wdm_sym(const std::string& path, uint64_t rows, uint64_t cols, uint32_t p,
                  const uint32_t* dr, const uint32_t* dc, const uint32_t* v, const uint32_t* curv,
                  uint32_t b, const uint32_t* seq, uint64_t L, unsigned threads) {
    if (L == 0) throw std::runtime_error("Cannot save WDM file with empty sequence list");
    if (threads == 0) threads = writer_threads();
    const std::string tmp = path + ".tmp";
    {
        TextSink s(tmp, path_requests_compression(tmp), threads);
        put_prefix(s, rows, cols, p, dr, dc, v, b, L, threads);
        for (uint32_t c = 0; c < b; ++c) put_line(s, curv + c, cols, b, p, threads);
        const uint64_t ut = (uint64_t)b * (b + 1) / 2;
        for (uint64_t e = 0; e < ut; ++e) put_line(s, seq + e, L, ut, p, threads);
        s.finish();
    }
    std::filesystem::rename(tmp, path);
}

struct WdmStream::Impl {
    std::string path, tmp;
    uint64_t rows = 0, cols = 0, L = 0, ut = 0;
    uint32_t p = 0, b = 0;
    unsigned threads = 1;
    const uint32_t *dr = nullptr, *dc = nullptr, *v = nullptr;
    bool buffer_lines = false;
    std::vector<std::string> lines;  // entry e's values of rows [0, formatted), space separated

    std::mutex mu;
    std::condition_variable cv;
    const uint32_t* seq = nullptr;  // guarded by mu
    const uint32_t* curv = nullptr;
    uint64_t ready = 0;             // rows of seq with final values
    bool finishing = false, stop = false;
    std::exception_ptr err;
    std::thread worker;

    // rows [t0, t1) of every entry appended to its line
    void format_rows(const uint32_t* s, uint64_t t0, uint64_t t1) {
        char tmp_c[16];
        for (uint64_t e = 0; e < ut; ++e) {
            std::string& out = lines[e];
            for (uint64_t t = t0; t < t1; ++t) {
                if (t) out.push_back(' ');
                auto r = std::to_chars(tmp_c, tmp_c + sizeof(tmp_c), s[t * ut + e] % p);
                out.append(tmp_c, (size_t)(r.ptr - tmp_c));
            }
        }
    }

    void run() {
        try {
            TextSink s(tmp, path_requests_compression(tmp), threads);
            put_prefix(s, rows, cols, p, dr, dc, v, b, L, threads);
            uint64_t formatted = 0;
            const uint32_t* sq = nullptr;
            const uint32_t* cv_block = nullptr;
            for (;;) {
                uint64_t upto;
                bool fin;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || finishing || (buffer_lines && ready > formatted); });
                    if (stop) return;
                    upto = ready;
                    fin = finishing;
                    sq = seq;
                    cv_block = curv;
                }
                if (buffer_lines && upto > formatted) {
                    format_rows(sq, formatted, upto);
                    formatted = upto;
                }
                if (fin) break;
            }
            for (uint32_t c = 0; c < b; ++c) put_line(s, cv_block + c, cols, b, p, threads);
            for (uint64_t e = 0; e < ut; ++e) {
                if (buffer_lines) {
                    s.put(lines[e].data(), lines[e].size());
                    s.put_char('\n');
                    std::string().swap(lines[e]);
                } else {
                    put_line(s, sq + e, L, ut, p, threads);
                }
            }
            s.finish();
            std::filesystem::rename(tmp, path);
        } catch (...) {
            err = std::current_exception();
        }
    }
};

WdmStream::WdmStream(const std::string& path, uint64_t rows, uint64_t cols, uint32_t p, const uint32_t* dr,
                     const uint32_t* dc, const uint32_t* v, uint32_t b, uint64_t L, unsigned threads,
                     uint64_t max_text_bytes)
    : impl_(new Impl) {
    if (L == 0) throw std::runtime_error("Cannot save WDM file with empty sequence list");
    Impl& m = *impl_;
    m.path = path;
    m.tmp = path + ".tmp";
    m.rows = rows;
    m.cols = cols;
    m.p = p;
    m.dr = dr;
    m.dc = dc;
    m.v = v;
    m.b = b;
    m.L = L;
    m.ut = (uint64_t)b * (b + 1) / 2;
    m.threads = threads ? threads : writer_threads();
    // at most 11 characters per value (10 digits and a separator)
    m.buffer_lines = L <= max_text_bytes / 11 / m.ut;
    if (m.buffer_lines) {
        m.lines.resize(m.ut);
        for (auto& s : m.lines) s.reserve(L * 11);
    }
    m.worker = std::thread([this] { impl_->run(); });
}

WdmStream::~WdmStream() {
    Impl& m = *impl_;
    if (m.worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(m.mu);
            m.stop = true;
        }
        m.cv.notify_all();
        m.worker.join();
        std::error_code ec;
        std::filesystem::remove(m.tmp, ec);
    }
}

void WdmStream::terms(const uint32_t* seq, uint64_t done) {
    Impl& m = *impl_;
    if (!seq) throw std::runtime_error("WDM stream: null sequence");
    {
        std::lock_guard<std::mutex> lk(m.mu);
        if (m.finishing) throw std::runtime_error("WDM stream: terms after finish");
        m.seq = seq;
        m.ready = std::max(m.ready, std::min(done, m.L));
    }
    m.cv.notify_all();
}

void WdmStream::finish(const uint32_t* curv) {
    Impl& m = *impl_;
    if (!m.worker.joinable()) throw std::runtime_error("WDM stream: already finished");
    {
        std::lock_guard<std::mutex> lk(m.mu);
        if (!curv || !m.seq) throw std::runtime_error("WDM stream: finish needs the sequence and curV");
        m.curv = curv;
        m.ready = m.L;
        m.finishing = true;
    }
    m.cv.notify_all();
    m.worker.join();
    if (m.err) {
        std::error_code ec;
        std::filesystem::remove(m.tmp, ec);
        std::rethrow_exception(m.err);
    }
}

namespace {
struct Scanner {
    const std::string& t;
    size_t pos = 0;
    explicit Scanner(const std::string& text) : t(text) {}
    uint64_t next(const char* what) {
        while (pos < t.size() && (t[pos] == ' ' || t[pos] == '\n' || t[pos] == '\r' || t[pos] == '\t')) ++pos;
        if (pos >= t.size() || t[pos] < '0' || t[pos] > '9')
            throw std::runtime_error(std::string("Failed to read ") + what + " from WDM file");
        uint64_t v = 0;
        while (pos < t.size() && t[pos] >= '0' && t[pos] <= '9') v = v * 10 + (uint64_t)(t[pos++] - '0');
        return v;
    }
};

void read_initial(Scanner& sc, WdmInitial& w) {
    w.rows = sc.next("matrix row count");
    w.cols = sc.next("matrix column count");
    w.prime = (uint32_t)sc.next("prime");
    w.length = sc.next("sequence length");
    w.block = (uint32_t)sc.next("vector count");
    w.dr.resize(w.rows);
    for (auto& x : w.dr) x = (uint32_t)sc.next("row preconditioner entry");
    w.dc.resize(w.cols);
    for (auto& x : w.dc) x = (uint32_t)sc.next("column preconditioner entry");
    w.v.resize(w.cols * w.block);
    for (uint32_t c = 0; c < w.block; ++c)
        for (uint64_t r = 0; r < w.cols; ++r) w.v[r * w.block + c] = (uint32_t)sc.next("initial vector entry");
}
}  // namespace

WdmInitial load_wdm_initial(const std::string& path) {
    std::string text = read_maybe_compressed(path);
    Scanner sc(text);
    WdmInitial w;
    read_initial(sc, w);
    return w;
}

WdmFull load_wdm_full(const std::string& path) {
    std::string text = read_maybe_compressed(path);
    Scanner sc(text);
    WdmFull w;
    read_initial(sc, w);
    w.curv.resize(w.cols * w.block);
    for (uint32_t c = 0; c < w.block; ++c)
        for (uint64_t r = 0; r < w.cols; ++r) w.curv[r * w.block + c] = (uint32_t)sc.next("current vector entry");
    const uint64_t ut = (uint64_t)w.block * (w.block + 1) / 2;
    w.seq.resize(w.length * ut);
    for (uint64_t e = 0; e < ut; ++e)
        for (uint64_t t = 0; t < w.length; ++t) w.seq[t * ut + e] = (uint32_t)sc.next("sequence entry");
    return w;
}

}  // namespace pgwio
