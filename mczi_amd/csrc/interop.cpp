// On-disk formats of the reference's partitioned MC path, for interop with
// it and for a spill mode: the packed-read cache (write_packed_read_cache_batch
// / for_packed_read_cache_batches, src/lib.rs:3394-3450), super-k-mer partition
// files (append_record_len 3209-3216, add_packed_range 4893-4917,
// read_superkmer_partition 2692-2711, read_record_len 2756-2769, file names
// 2827-2833) and counted k-mer partitions (process_superkmer_partition
// 2413-2421, emit_counted_partition 2657-2678). All integers are big-endian;
// packed bases use packed-seq codes (A0 C1 T2 G3), 4 per byte from bit 0,
// unused high bits of a last byte zero.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"

namespace mczi {

namespace {

const char kCacheMagic[8] = {'M', 'C', 'R', 'D', '0', '0', '0', '1'};

std::string read_whole(const std::string &path) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) throw Error(MCZI_E_IO, "failed to open " + path + ": " + strerror(errno));
    std::string s;
    char buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) throw Error(MCZI_E_IO, "failed to read " + path);
    return s;
}

void write_whole(const std::string &path, const std::string &s) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) throw Error(MCZI_E_IO, "failed to create " + path + ": " + strerror(errno));
    const bool ok = fwrite(s.data(), 1, s.size(), f) == s.size();
    if (fclose(f) != 0 || !ok) throw Error(MCZI_E_IO, "failed to write " + path);
}

void put_be(std::string &s, uint64_t v, int bytes) {
    for (int i = bytes - 1; i >= 0; i--) s.push_back((char)(uint8_t)(v >> (8 * i)));
}

uint64_t get_be(const std::string &s, size_t at, int bytes) {
    uint64_t v = 0;
    for (int i = 0; i < bytes; i++) v = (v << 8) | (uint8_t)s[at + i];
    return v;
}

// Append len packed bases (4 per byte) as one fragment of out, dropping it
// when shorter than min_len.
void append_fragment(HostReads &out, const uint8_t *bytes, uint64_t len, uint64_t min_len, uint32_t ds) {
    if (len == 0 || len < min_len) return;
    const uint64_t p0 = out.cur_;
    out.words.resize((p0 + len + 15) / 16 + 1, 0);
    uint32_t *w = out.words.data();
    const uint64_t nb = (len + 3) / 4;
    for (uint64_t j = 0; j < nb; j++) {
        uint32_t b = bytes[j];
        const uint64_t rem = len - 4 * j;
        if (rem < 4) b &= (1u << (2 * rem)) - 1u;
        const uint64_t p = p0 + 4 * j;
        const unsigned sh = (unsigned)(p & 15u) * 2u;
        w[p >> 4] |= b << sh;
        if (sh > 24) w[(p >> 4) + 1] |= b >> (32u - sh);
    }
    out.cur_ = p0 + len;
    out.words.resize((out.cur_ + 15) / 16);
    out.off.push_back(out.cur_);
    out.dataset.push_back(ds);
}

// append_record_len (src/lib.rs:3209-3216)
void put_record_len(std::string &s, uint32_t len) {
    if (len < 255) {
        s.push_back((char)(uint8_t)len);
    } else {
        s.push_back((char)0xFF);
        put_be(s, len, 4);
    }
}

}  // namespace

void write_read_cache(const std::string &path, const HostReads &r) {
    std::string s(kCacheMagic, 8);
    const uint32_t *w = r.words.data();
    for (uint64_t f = 0; f + 1 < r.off.size(); f++) {
        const uint64_t a = r.off[f], len = r.off[f + 1] - a;
        require(len <= 0xFFFFFFFFull, MCZI_E_INVALID, "read is too large for packed-read cache format");
        put_be(s, len, 4);
        for (uint64_t j = 0; j < len; j += 4) {
            uint32_t b = 0;
            for (uint64_t i = j; i < std::min(len, j + 4); i++) {
                const uint64_t p = a + i;
                b |= ((w[p >> 4] >> ((p & 15u) * 2u)) & 3u) << (2 * (i - j));
            }
            s.push_back((char)(uint8_t)b);
        }
    }
    write_whole(path, s);
}

void read_read_cache(const std::string &path, uint64_t min_len, HostReads &out) {
    const std::string s = read_whole(path);
    require(s.size() >= 8 && memcmp(s.data(), kCacheMagic, 8) == 0, MCZI_E_INVALID,
            "invalid packed-read cache header in " + path);
    size_t at = 8;
    while (at < s.size()) {
        require(at + 4 <= s.size(), MCZI_E_INVALID, "truncated packed-read cache record in " + path);
        const uint64_t len = get_be(s, at, 4);
        at += 4;
        const uint64_t nb = (len + 3) / 4;
        require(at + nb <= s.size(), MCZI_E_INVALID, "truncated packed-read cache record in " + path);
        append_fragment(out, (const uint8_t *)s.data() + at, len, min_len, 0);
        at += nb;
    }
    out.end_chunk();
}

void write_superkmer_partitions(const std::string &dir, uint32_t partitions, const std::vector<uint64_t> &rec,
                                const std::vector<uint32_t> &part, uint32_t rec_words) {
    require(partitions > 0, MCZI_E_INVALID, "partition count must be greater than 0");
    const size_t n = part.size();
    // records of each user partition, keeping their device order
    std::vector<size_t> start(partitions + 1, 0);
    for (size_t i = 0; i < n; i++) start[part[i] % partitions + 1]++;
    for (uint32_t u = 0; u < partitions; u++) start[u + 1] += start[u];
    std::vector<uint32_t> order(n);
    {
        std::vector<size_t> at(start.begin(), start.end() - 1);
        for (size_t i = 0; i < n; i++) order[at[part[i] % partitions]++] = (uint32_t)i;
    }
    std::vector<std::string> errors(partitions);
    parallel_ranges(partitions, 1, host_workers((size_t)1 << 20), [&](size_t u0, size_t u1) {
        for (size_t u = u0; u < u1; u++) {
            try {
                std::string s;
                for (size_t o = start[u]; o < start[u + 1]; o++) {
                    const uint64_t *r = rec.data() + (size_t)order[o] * rec_words;
                    const uint32_t len = (uint32_t)(r[rec_words - 1] >> 48) & 255u;
                    put_record_len(s, len);
                    for (uint32_t j = 0; j < len; j += 4) {
                        uint32_t b = 0;
                        for (uint32_t i = j; i < std::min(len, j + 4); i++) {
                            const uint32_t mc = (uint32_t)(r[i >> 5] >> (2 * (i & 31))) & 3u;
                            b |= (mc ^ (mc >> 1)) << (2 * (i - j));  // MC code -> packed-seq code
                        }
                        s.push_back((char)(uint8_t)b);
                    }
                }
                char name[32];
                snprintf(name, sizeof name, "/partition_%04zu.skm", u);
                write_whole(dir + name, s);
            } catch (const std::exception &e) {
                errors[u] = e.what();
            }
        }
    });
    for (const auto &e : errors) require(e.empty(), MCZI_E_IO, e);
}

void read_superkmer_file(const std::string &path, uint64_t min_len, HostReads &out) {
    const std::string s = read_whole(path);
    size_t at = 0;
    while (at < s.size()) {
        uint64_t len = (uint8_t)s[at++];
        if (len == 255) {
            require(at + 4 <= s.size(), MCZI_E_INVALID, "truncated super-kmer record in " + path);
            len = get_be(s, at, 4);
            at += 4;
        }
        const uint64_t nb = (len + 3) / 4;
        require(at + nb <= s.size(), MCZI_E_INVALID, "truncated super-kmer record in " + path);
        append_fragment(out, (const uint8_t *)s.data() + at, len, min_len, 0);
        at += nb;
    }
    out.end_chunk();
}

void write_counted_file(const std::string &path, const mczi_counted_kmer *kmers, size_t n) {
    std::string s;
    s.reserve(n * 24);
    for (size_t i = 0; i < n; i++) {
        put_be(s, kmers[i].hi, 8);
        put_be(s, kmers[i].lo, 8);
        put_be(s, kmers[i].count, 8);
    }
    write_whole(path, s);
}

std::vector<mczi_counted_kmer> read_counted_file(const std::string &path) {
    const std::string s = read_whole(path);
    require(s.size() % 24 == 0, MCZI_E_INVALID, "truncated counted k-mer record in " + path);
    std::vector<mczi_counted_kmer> v(s.size() / 24);
    for (size_t i = 0; i < v.size(); i++)
        v[i] = mczi_counted_kmer{get_be(s, 24 * i + 8, 8), get_be(s, 24 * i, 8), get_be(s, 24 * i + 16, 8)};
    return v;
}

}  // namespace mczi
