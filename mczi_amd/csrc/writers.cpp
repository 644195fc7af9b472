// Host end of the path: simplitig FASTA, per-k-mer FASTA and KFF writers.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "internal.hpp"

namespace mczi {

namespace {

typedef unsigned __int128 u128;

inline u128 enc_of(const mczi_counted_kmer &r) { return ((u128)r.hi << 64) | r.lo; }

const char kAscii[4] = {'A', 'C', 'G', 'T'};

class Out {
  public:
    // .gz / .gzip, .xz and .zst / .zstd paths are written compressed
    explicit Out(const std::string &path) : path_(path) {
        fp_ = fopen(path.c_str(), "wb");
        if (!fp_) throw Error(MCZI_E_IO, "failed to create " + path);
        if (const int kind = output_compression_kind(path)) z_.reset(new Compressor(fp_, path, kind));
        buf_.reserve(1 << 22);
    }
    explicit Out(std::string *sink) : sink_(sink) {}
    ~Out() {
        z_.reset();
        if (fp_) fclose(fp_);
    }
    void put(const char *s, size_t n) {
        if (sink_) {
            sink_->append(s, n);
            return;
        }
        buf_.append(s, n);
        if (buf_.size() >= (1u << 22)) flush();
    }
    void put(const std::string &s) { put(s.data(), s.size()); }
    void flush(bool last = false) {
        if (!fp_) return;
        if (z_) {
            if (!buf_.empty() || last) z_->write(buf_.data(), buf_.size(), last);
            buf_.clear();
            return;
        }
        if (!buf_.empty()) {
            if (fwrite(buf_.data(), 1, buf_.size(), fp_) != buf_.size())
                throw Error(MCZI_E_IO, "failed to write " + path_);
            buf_.clear();
        }
    }
    void finish() {
        flush(true);
        z_.reset();
        if (fp_) {
            if (fclose(fp_) != 0) {
                fp_ = nullptr;
                throw Error(MCZI_E_IO, "failed to flush " + path_);
            }
            fp_ = nullptr;
        }
    }

  private:
    std::string path_;
    FILE *fp_ = nullptr;
    std::string *sink_ = nullptr;
    std::string buf_;
    std::unique_ptr<Compressor> z_;
};

void put_lines(Out &o, const std::vector<uint8_t> &seq) {
    char line[81];
    for (size_t i = 0; i < seq.size(); i += 80) {
        size_t c = std::min<size_t>(80, seq.size() - i);
        for (size_t j = 0; j < c; j++) line[j] = kAscii[seq[i + j]];
        line[c] = '\n';
        o.put(line, c + 1);
    }
}

}  // namespace

void write_simplitig_fasta(const std::string &path, const mczi_counted_kmer *kmers, size_t n,
                           uint32_t k, size_t first_index, size_t *last_index) {
    Out o(path);
    const size_t last =
        emit_simplitigs(kmers, n, k, first_index, [&](const char *t, size_t len) { o.put(t, len); });
    if (last_index) *last_index = last;
    o.finish();
}

void write_simplitig_fasta_keys(const std::string &path, size_t n, uint32_t k,
                                const std::function<void(uint64_t *, uint64_t *)> &fill,
                                const std::function<void(uint64_t *)> &link_fill) {
    Out o(path);
    emit_simplitigs_keys(n, k, fill, 0, [&](const char *t, size_t len) { o.put(t, len); }, link_fill);
    o.finish();
}

std::string simplitig_fasta_text(const mczi_counted_kmer *kmers, size_t n, uint32_t k) {
    std::string s;
    Out o(&s);
    emit_simplitigs(kmers, n, k, 0, [&](const char *t, size_t len) { o.put(t, len); });
    return s;
}

// write_kmer_fasta_record_u128 (src/lib.rs:1535-1544)
void write_kmer_fasta(const std::string &path, const std::vector<mczi_counted_kmer> &kmers,
                      uint32_t k) {
    Out o(path);
    std::vector<uint8_t> seq(k);
    char head[96];
    for (size_t i = 0; i < kmers.size(); i++) {
        const u128 e = enc_of(kmers[i]);
        for (uint32_t j = 0; j < k; j++) seq[j] = (uint8_t)((e >> (2 * (k - 1 - j))) & 3u);
        int hl = snprintf(head, sizeof head, ">MC_kmer_%zu count=%llu\n", i + 1,
                          (unsigned long long)kmers[i].count);
        o.put(head, (size_t)hl);
        put_lines(o, seq);
    }
    o.finish();
}

// write_kff_header_with_count_size (src/lib.rs:1633-1664) + records
// (write_packed_dna 3082-3087, write_kff_u8_record_u64 3089-3102,
// write_kff_u32_record_u64 3104-3117) + "KFF" footer.
void write_kff(const std::string &path, const std::vector<mczi_counted_kmer> &kmers,
               const mczi_counter_config &cfg, uint64_t count_size, bool ordered,
               uint64_t saturate_at) {
    Out o(path);
    std::string h = "KFF";
    h.push_back(1);
    h.push_back(0);
    h.push_back(0x1b);
    h.push_back(1);
    h.push_back(1);
    char fb[256];
    int fl = snprintf(fb, sizeof fb,
                      "producer=MC\nk=%u\nm=%u\nthreshold=%llu\nfilter=count>threshold\ncanonical=1\n",
                      cfg.k, cfg.minimizer, (unsigned long long)cfg.threshold);
    auto be32 = [&](uint32_t v) {
        for (int i = 3; i >= 0; i--) h.push_back((char)(v >> (8 * i)));
    };
    auto be64 = [&](uint64_t v) {
        for (int i = 7; i >= 0; i--) h.push_back((char)(v >> (8 * i)));
    };
    be32((uint32_t)fl);
    h.append(fb, (size_t)fl);
    h.push_back('v');
    be64(6);
    const char *names[6] = {"k", "max", "data_size", "ordered", "m", "threshold"};
    const uint64_t vals[6] = {cfg.k, 1, count_size, ordered ? 1u : 0u, cfg.minimizer, cfg.threshold};
    for (int i = 0; i < 6; i++) {
        h.append(names[i]);
        h.push_back('\0');
        be64(vals[i]);
    }
    h.push_back('r');
    be64(kmers.size());
    o.put(h);
    const size_t bytes = (cfg.k + 3) / 4;
    std::string rec;
    rec.reserve(1 << 20);
    for (const auto &r : kmers) {
        const u128 e = enc_of(r);
        for (size_t i = 0; i < bytes; i++) rec.push_back((char)(uint8_t)(e >> (8 * (bytes - 1 - i))));
        uint64_t c = r.count;
        if (saturate_at && c > saturate_at) c = saturate_at;
        for (int i = (int)count_size - 1; i >= 0; i--) rec.push_back((char)(uint8_t)(c >> (8 * i)));
        if (rec.size() >= (1u << 20)) {
            o.put(rec);
            rec.clear();
        }
    }
    rec.append("KFF");
    o.put(rec);
    o.finish();
}

}  // namespace mczi
