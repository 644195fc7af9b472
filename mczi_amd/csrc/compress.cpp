// Compressed inputs and outputs: gzip through zlib, xz and zstd through the
// system's liblzma.so.5 / libzstd.so.1, loaded on first use (the image ships
// their runtime libraries, not their headers, so the few stable public entry
// points used here are declared below).
//
// Inputs are recognised by their magic bytes, as FASTX readers that accept
// compressed files do (src/lib.rs:3218-3450 reads .gz / .xz / .zst inputs;
// with_xz_decompressed_path, src/lib.rs:1407-1447); outputs by the path's
// extension, gz / gzip, xz, zst / zstd, case-insensitive
// (output_compression, src/lib.rs:1383-1395), at the reference's levels:
// gzip fast (1), xz 6, zstd 1 (create_output_writer, src/lib.rs:1357-1381).
#include <dlfcn.h>
#include <zlib.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "internal.hpp"

namespace mczi {

namespace {

// ---- liblzma (xz) --------------------------------------------------------
struct LzmaStream {  // lzma_stream of liblzma 5.x
    const uint8_t *next_in;
    size_t avail_in;
    uint64_t total_in;
    uint8_t *next_out;
    size_t avail_out;
    uint64_t total_out;
    const void *allocator;
    void *internal;
    void *reserved_ptr[4];
    uint64_t reserved_int[2];
    size_t reserved_sz[2];
    int reserved_enum[2];
    uint64_t tail_guard[8];  // zero; never touched by the library
};
constexpr int LZMA_OK_ = 0, LZMA_STREAM_END_ = 1;
constexpr int LZMA_RUN_ = 0, LZMA_FINISH_ = 3;
constexpr uint32_t LZMA_CONCATENATED_ = 0x08;
constexpr int LZMA_CHECK_CRC64_ = 4;

struct LzmaApi {
    int (*stream_decoder)(LzmaStream *, uint64_t, uint32_t) = nullptr;
    int (*easy_encoder)(LzmaStream *, uint32_t, int) = nullptr;
    int (*code)(LzmaStream *, int) = nullptr;
    void (*end)(LzmaStream *) = nullptr;
};

// ---- libzstd -----------------------------------------------------------------
struct ZInBuf {
    const void *src;
    size_t size;
    size_t pos;
};
struct ZOutBuf {
    void *dst;
    size_t size;
    size_t pos;
};
struct ZstdApi {
    void *(*create_d)() = nullptr;
    size_t (*init_d)(void *) = nullptr;
    size_t (*decompress_stream)(void *, ZOutBuf *, ZInBuf *) = nullptr;
    size_t (*free_d)(void *) = nullptr;
    void *(*create_c)() = nullptr;
    size_t (*init_c)(void *, int) = nullptr;
    size_t (*compress_stream)(void *, ZOutBuf *, ZInBuf *) = nullptr;
    size_t (*end_stream)(void *, ZOutBuf *) = nullptr;
    size_t (*free_c)(void *) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    const char *(*error_name)(size_t) = nullptr;
};

template <class F>
void bind(void *h, const char *name, F &f, const char *lib) {
    f = reinterpret_cast<F>(dlsym(h, name));
    if (!f) throw Error(MCZI_E_IO, std::string(lib) + " lacks " + name);
}

const LzmaApi &lzma() {
    static LzmaApi api;
    static std::once_flag once;
    static std::string err;
    std::call_once(once, [] {
        void *h = dlopen("liblzma.so.5", RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            err = "xz support needs liblzma.so.5, which could not be loaded";
            return;
        }
        try {
            bind(h, "lzma_stream_decoder", api.stream_decoder, "liblzma");
            bind(h, "lzma_easy_encoder", api.easy_encoder, "liblzma");
            bind(h, "lzma_code", api.code, "liblzma");
            bind(h, "lzma_end", api.end, "liblzma");
        } catch (const Error &e) {
            err = e.what();
            api = LzmaApi();
        }
    });
    if (!api.code) throw Error(MCZI_E_IO, err);
    return api;
}

const ZstdApi &zstd() {
    static ZstdApi api;
    static std::once_flag once;
    static std::string err;
    std::call_once(once, [] {
        void *h = dlopen("libzstd.so.1", RTLD_NOW | RTLD_LOCAL);
        if (!h) {
            err = "zstd support needs libzstd.so.1, which could not be loaded";
            return;
        }
        try {
            bind(h, "ZSTD_createDStream", api.create_d, "libzstd");
            bind(h, "ZSTD_initDStream", api.init_d, "libzstd");
            bind(h, "ZSTD_decompressStream", api.decompress_stream, "libzstd");
            bind(h, "ZSTD_freeDStream", api.free_d, "libzstd");
            bind(h, "ZSTD_createCStream", api.create_c, "libzstd");
            bind(h, "ZSTD_initCStream", api.init_c, "libzstd");
            bind(h, "ZSTD_compressStream", api.compress_stream, "libzstd");
            bind(h, "ZSTD_endStream", api.end_stream, "libzstd");
            bind(h, "ZSTD_freeCStream", api.free_c, "libzstd");
            bind(h, "ZSTD_isError", api.is_error, "libzstd");
            bind(h, "ZSTD_getErrorName", api.error_name, "libzstd");
        } catch (const Error &e) {
            err = e.what();
            api = ZstdApi();
        }
    });
    if (!api.decompress_stream) throw Error(MCZI_E_IO, err);
    return api;
}

// Growing output buffer: doubles, starting from a size hint.
struct Grow {
    std::string &s;
    size_t used = 0;
    Grow(std::string &s_, size_t hint) : s(s_) { s.resize(std::max<size_t>(hint, 1 << 20)); }
    char *room(size_t &n) {
        if (s.size() - used < (1 << 16)) s.resize(s.size() * 2);
        n = s.size() - used;
        return &s[used];
    }
    void done() { s.resize(used); }
};

void gunzip(const uint8_t *in, size_t len, const std::string &what, std::string &out) {
    // ISIZE (the last 4 bytes) is the size of the last member mod 2^32
    uint64_t hint = 4 * (uint64_t)len;
    if (len >= 4) {
        const uint32_t isz = (uint32_t)in[len - 4] | ((uint32_t)in[len - 3] << 8) | ((uint32_t)in[len - 2] << 16) |
                             ((uint32_t)in[len - 1] << 24);
        hint = std::max<uint64_t>(hint, isz);
    }
    Grow g(out, hint);
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit2(&zs, 15 + 32) != Z_OK) throw Error(MCZI_E_IO, "zlib init failed for " + what);
    size_t pos = 0;
    for (;;) {
        const size_t take = std::min<size_t>(len - pos, 1u << 30);
        zs.next_in = const_cast<Bytef *>(in + pos);
        zs.avail_in = (uInt)take;
        size_t n;
        char *dst = g.room(n);
        n = std::min<size_t>(n, 1u << 30);
        zs.next_out = reinterpret_cast<Bytef *>(dst);
        zs.avail_out = (uInt)n;
        const int rc = inflate(&zs, Z_NO_FLUSH);
        g.used += n - zs.avail_out;
        pos += take - zs.avail_in;
        if (rc == Z_STREAM_END) {
            // concatenated members (multi-member gzip) continue the stream
            while (pos < len && in[pos] == 0) pos++;  // trailing zero padding
            if (pos >= len) break;
            if (inflateReset(&zs) != Z_OK) break;
            continue;
        }
        if (rc == Z_BUF_ERROR && zs.avail_in == 0 && pos >= len) {
            inflateEnd(&zs);
            throw Error(MCZI_E_IO, "truncated gzip input " + what);
        }
        if (rc != Z_OK && rc != Z_BUF_ERROR) {
            inflateEnd(&zs);
            throw Error(MCZI_E_IO, "corrupt gzip input " + what + (zs.msg ? std::string(": ") + zs.msg : ""));
        }
    }
    inflateEnd(&zs);
    g.done();
}

void unxz(const uint8_t *in, size_t len, const std::string &what, std::string &out) {
    const LzmaApi &L = lzma();
    LzmaStream st;
    memset(&st, 0, sizeof st);
    if (L.stream_decoder(&st, UINT64_MAX, LZMA_CONCATENATED_) != LZMA_OK_)
        throw Error(MCZI_E_IO, "xz decoder init failed for " + what);
    Grow g(out, 6 * (uint64_t)len);
    st.next_in = in;
    st.avail_in = len;
    for (;;) {
        size_t n;
        char *dst = g.room(n);
        st.next_out = reinterpret_cast<uint8_t *>(dst);
        st.avail_out = n;
        const int rc = L.code(&st, LZMA_FINISH_);
        g.used += n - st.avail_out;
        if (rc == LZMA_STREAM_END_) break;
        if (rc != LZMA_OK_) {
            L.end(&st);
            throw Error(MCZI_E_IO, "corrupt xz input " + what + " (liblzma code " + std::to_string(rc) + ")");
        }
    }
    L.end(&st);
    g.done();
}

void unzstd(const uint8_t *in, size_t len, const std::string &what, std::string &out) {
    const ZstdApi &Z = zstd();
    void *ds = Z.create_d();
    if (!ds) throw Error(MCZI_E_IO, "zstd decoder init failed for " + what);
    Z.init_d(ds);
    Grow g(out, 4 * (uint64_t)len);
    ZInBuf ib{in, len, 0};
    size_t last = 0;
    while (ib.pos < ib.size || last != 0) {
        size_t n;
        char *dst = g.room(n);
        ZOutBuf ob{dst, n, 0};
        const size_t rc = Z.decompress_stream(ds, &ob, &ib);
        g.used += ob.pos;
        if (Z.is_error(rc)) {
            const std::string e = Z.error_name(rc);
            Z.free_d(ds);
            throw Error(MCZI_E_IO, "corrupt zstd input " + what + ": " + e);
        }
        last = rc;
        if (ib.pos >= ib.size && rc != 0 && ob.pos == 0) {
            Z.free_d(ds);
            throw Error(MCZI_E_IO, "truncated zstd input " + what);
        }
    }
    Z.free_d(ds);
    g.done();
}

std::string lower_ext(const std::string &path) {
    const size_t slash = path.find_last_of('/');
    const size_t dot = path.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash)) return "";
    std::string e = path.substr(dot + 1);
    for (char &c : e) c = (char)tolower((unsigned char)c);
    return e;
}

}  // namespace

InputCompression input_compression(const uint8_t *p, size_t len) {
    if (len >= 2 && p[0] == 0x1f && p[1] == 0x8b) return InputCompression::Gzip;
    if (len >= 6 && p[0] == 0xfd && p[1] == '7' && p[2] == 'z' && p[3] == 'X' && p[4] == 'Z' && p[5] == 0)
        return InputCompression::Xz;
    if (len >= 4 && p[0] == 0x28 && p[1] == 0xb5 && p[2] == 0x2f && p[3] == 0xfd) return InputCompression::Zstd;
    return InputCompression::None;
}

void decompress_input(const uint8_t *p, size_t len, InputCompression c, const std::string &what, std::string &out) {
    switch (c) {
        case InputCompression::Gzip: gunzip(p, len, what, out); break;
        case InputCompression::Xz: unxz(p, len, what, out); break;
        case InputCompression::Zstd: unzstd(p, len, what, out); break;
        case InputCompression::None: out.assign((const char *)p, len); break;
    }
}

// ---- compressed output -------------------------------------------------------
struct Compressor::Impl {
    int kind = 0;  // 1 gzip, 2 xz, 3 zstd
    FILE *fp = nullptr;
    std::string path;
    std::string obuf;
    z_stream zs;
    LzmaStream xs;
    void *zc = nullptr;

    void put_out(const char *p, size_t n) {
        if (n && fwrite(p, 1, n, fp) != n) throw Error(MCZI_E_IO, "failed to write " + path);
    }
};

int output_compression_kind(const std::string &path) {
    const std::string e = lower_ext(path);
    if (e == "gz" || e == "gzip") return 1;
    if (e == "xz") return 2;
    if (e == "zst" || e == "zstd") return 3;
    return 0;
}

Compressor::Compressor(FILE *fp, const std::string &path, int kind) : m_(new Impl) {
    m_->kind = kind;
    m_->fp = fp;
    m_->path = path;
    m_->obuf.resize(1 << 20);
    if (kind == 1) {
        memset(&m_->zs, 0, sizeof m_->zs);
        if (deflateInit2(&m_->zs, 1, Z_DEFLATED, 15 + 16, 8, Z_DEFAULT_STRATEGY) != Z_OK)
            throw Error(MCZI_E_IO, "zlib init failed for " + path);
    } else if (kind == 2) {
        memset(&m_->xs, 0, sizeof m_->xs);
        if (lzma().easy_encoder(&m_->xs, 6, LZMA_CHECK_CRC64_) != LZMA_OK_)
            throw Error(MCZI_E_IO, "xz encoder init failed for " + path);
    } else if (kind == 3) {
        const ZstdApi &Z = zstd();
        m_->zc = Z.create_c();
        if (!m_->zc || Z.is_error(Z.init_c(m_->zc, 1))) throw Error(MCZI_E_IO, "zstd encoder init failed for " + path);
    }
}

Compressor::~Compressor() {
    if (!m_) return;
    if (m_->kind == 1) deflateEnd(&m_->zs);
    if (m_->kind == 2 && m_->xs.internal) lzma().end(&m_->xs);
    if (m_->kind == 3 && m_->zc) zstd().free_c(m_->zc);
    delete m_;
}

void Compressor::write(const char *p, size_t n, bool finish) {
    Impl &m = *m_;
    char *ob = &m.obuf[0];
    const size_t on = m.obuf.size();
    if (m.kind == 1) {
        size_t pos = 0;
        for (;;) {
            const size_t take = std::min<size_t>(n - pos, 1u << 30);
            m.zs.next_in = reinterpret_cast<Bytef *>(const_cast<char *>(p + pos));
            m.zs.avail_in = (uInt)take;
            const bool last = finish && pos + take == n;
            int rc;
            do {
                m.zs.next_out = reinterpret_cast<Bytef *>(ob);
                m.zs.avail_out = (uInt)on;
                rc = deflate(&m.zs, last ? Z_FINISH : Z_NO_FLUSH);
                if (rc == Z_STREAM_ERROR) throw Error(MCZI_E_IO, "gzip compression failed for " + m.path);
                m.put_out(ob, on - m.zs.avail_out);
            } while (m.zs.avail_out == 0 || (last && rc != Z_STREAM_END));
            pos += take;
            if (pos >= n) break;
        }
    } else if (m.kind == 2) {
        const LzmaApi &L = lzma();
        m.xs.next_in = reinterpret_cast<const uint8_t *>(p);
        m.xs.avail_in = n;
        for (;;) {
            m.xs.next_out = reinterpret_cast<uint8_t *>(ob);
            m.xs.avail_out = on;
            const int rc = L.code(&m.xs, finish ? LZMA_FINISH_ : LZMA_RUN_);
            m.put_out(ob, on - m.xs.avail_out);
            if (rc == LZMA_STREAM_END_) break;
            if (rc != LZMA_OK_) throw Error(MCZI_E_IO, "xz compression failed for " + m.path);
            if (!finish && m.xs.avail_in == 0 && m.xs.avail_out != 0) break;
        }
    } else if (m.kind == 3) {
        const ZstdApi &Z = zstd();
        ZInBuf ib{p, n, 0};
        while (ib.pos < ib.size) {
            ZOutBuf o{ob, on, 0};
            const size_t rc = Z.compress_stream(m.zc, &o, &ib);
            if (Z.is_error(rc)) throw Error(MCZI_E_IO, "zstd compression failed for " + m.path);
            m.put_out(ob, o.pos);
        }
        if (finish) {
            size_t rem;
            do {
                ZOutBuf o{ob, on, 0};
                rem = Z.end_stream(m.zc, &o);
                if (Z.is_error(rem)) throw Error(MCZI_E_IO, "zstd compression failed for " + m.path);
                m.put_out(ob, o.pos);
            } while (rem != 0);
        }
    }
}

}  // namespace mczi
