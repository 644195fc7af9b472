// Seeded synthetic genomes and reads (test and benchmark inputs).
//
// PRNG: SplitMix64 (bench/simulate_long_reads.rs:9-30); genome bases drawn
// 2 bits at a time from successive outputs (fill_random_genome,
// bench/simulate_long_reads.rs:166-180, bases A,C,G,T for values 0..3).
// Extensions for the workload shapes: random strand, iid per-base
// substitution / insertion / deletion, lognormal long reads, repeat copies,
// per-dataset SNP variants. Read i uses its own generator seeded from
// (read_seed, i), so output does not depend on thread count.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>

#include "internal.hpp"

namespace mczi {
namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint64_t range(uint64_t n) { return next() % n; }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// ACGT index (0..3) -> packed-seq code (A0 C1 T2 G3)
const uint8_t kAcgtToCode[4] = {0, 1, 3, 2};

// genome as packed-seq codes, one byte per base
std::vector<uint8_t> make_genome(uint64_t len, uint64_t seed) {
    std::vector<uint8_t> g(len);
    SplitMix64 rng(seed);
    uint64_t bits = 0;
    int avail = 0;
    for (uint64_t i = 0; i < len; i++) {
        if (avail < 2) {
            bits = rng.next();
            avail = 64;
        }
        g[i] = kAcgtToCode[bits & 3u];
        bits >>= 2;
        avail -= 2;
    }
    return g;
}

uint64_t read_seed(uint64_t base, uint64_t i) {
    SplitMix64 r(base ^ (i * 0xD1B54A32D192ED03ull));
    return r.next();
}

struct ReadSpec {
    uint64_t len;
    uint32_t dataset;
};

void sample_read(const std::vector<uint8_t> &genome, uint64_t len, double sub, double ins, double del,
                 SplitMix64 &rng, std::vector<uint8_t> &out) {
    out.clear();
    const uint64_t G = genome.size();
    if (len > G) len = G;
    // template span with room for deletions
    const uint64_t span = std::min<uint64_t>(G, len + len / 4 + 16);
    const uint64_t pos = rng.range(G - span + 1);
    const bool rc = rng.next() & 1;
    uint64_t t = 0;
    while (out.size() < len && t < span) {
        uint8_t c = rc ? (uint8_t)(genome[pos + span - 1 - t] ^ 2u)  // complement: code ^ 2
                       : genome[pos + t];
        const double u = rng.unit();
        if (u < sub) {
            out.push_back((uint8_t)((c + 1 + rng.range(3)) & 3u));
            t++;
        } else if (u < sub + ins) {
            out.push_back((uint8_t)rng.range(4));
        } else if (u < sub + ins + del) {
            t++;
        } else {
            out.push_back(c);
            t++;
        }
    }
}

}  // namespace

void synth_generate(const mczi_synth_spec &spec, mczi_synth_reads &out) {
    require(spec.genome_len > 0, MCZI_E_INVALID, "genome_len must be > 0");
    std::vector<uint8_t> base = make_genome(spec.genome_len, spec.genome_seed);
    // repeats: segments copied (repeat_copies - 1) extra times, appended
    if (spec.repeat_fraction > 0 && spec.repeat_len > 0 && spec.repeat_copies > 1) {
        SplitMix64 rr(spec.genome_seed ^ 0x5245504541545321ull);
        const uint64_t nseg = (uint64_t)(spec.genome_len * spec.repeat_fraction) / spec.repeat_len;
        std::vector<uint8_t> extra;
        for (uint64_t s = 0; s < nseg; s++) {
            const uint64_t p = rr.range(spec.genome_len - spec.repeat_len + 1);
            for (uint32_t c = 1; c < spec.repeat_copies; c++)
                extra.insert(extra.end(), base.begin() + p, base.begin() + p + spec.repeat_len);
        }
        base.insert(base.end(), extra.begin(), extra.end());
    }
    const uint32_t D = std::max<uint32_t>(spec.n_datasets, 1);
    std::vector<std::vector<uint8_t>> genomes;
    if (spec.n_datasets > 0) {
        for (uint32_t d = 0; d < D; d++) {
            std::vector<uint8_t> g = base;
            SplitMix64 sr(spec.genome_seed ^ (0x534E5000ull + d) * 0x9E3779B97F4A7C15ull);
            const uint64_t nsnp = (uint64_t)std::llround(g.size() * spec.snp_rate);
            for (uint64_t i = 0; i < nsnp; i++) {
                const uint64_t p = sr.range(g.size());
                g[p] = (uint8_t)((g[p] + 1 + sr.range(3)) & 3u);
            }
            genomes.push_back(std::move(g));
        }
    } else {
        genomes.push_back(std::move(base));
    }
    // read lengths
    std::vector<ReadSpec> reads;
    for (uint32_t d = 0; d < D; d++) {
        for (uint64_t i = 0; i < spec.n_reads; i++) reads.push_back({spec.read_len, d});
        if (spec.long_bases > 0) {
            SplitMix64 lr(spec.read_seed ^ 0x4C4F4E47ull ^ ((uint64_t)d << 40));
            uint64_t acc = 0;
            const double mu = std::log(spec.long_median > 0 ? spec.long_median : 8000.0);
            while (acc < spec.long_bases) {
                // Box-Muller normal
                double u1 = lr.unit(), u2 = lr.unit();
                if (u1 < 1e-300) u1 = 1e-300;
                const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                double L = std::exp(mu + spec.long_sigma * z);
                L = std::min<double>(std::max<double>(L, spec.long_min), spec.long_max);
                reads.push_back({(uint64_t)L, d});
                acc += (uint64_t)L;
            }
        }
    }
    // generate in parallel, each read from its own seed (contiguous read
    // ranges per thread, bases kept in a per-thread buffer)
    const size_t nr = reads.size();
    const unsigned nt = std::max(1u, host_workers(size_t(1) << 20));
    std::vector<std::vector<uint8_t>> tbuf(nt);
    std::vector<uint64_t> rlen(nr);
    auto rng_of = [&](unsigned t, size_t &b, size_t &e) {
        b = nr * t / nt;
        e = nr * (t + 1) / nt;
    };
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) {
            th.emplace_back([&, t]() {
                size_t b, e;
                rng_of(t, b, e);
                std::vector<uint8_t> tmp;
                for (size_t i = b; i < e; i++) {
                    SplitMix64 rng(read_seed(spec.read_seed, i));
                    sample_read(genomes[spec.n_datasets > 0 ? reads[i].dataset : 0], reads[i].len, spec.sub_rate,
                                spec.ins_rate, spec.del_rate, rng, tmp);
                    rlen[i] = tmp.size();
                    tbuf[t].insert(tbuf[t].end(), tmp.begin(), tmp.end());
                }
            });
        }
        for (auto &x : th) x.join();
    }
    out.n_frag = nr;
    out.frag_off = (uint64_t *)malloc((nr + 1) * 8);
    out.dataset = (uint32_t *)malloc(std::max<size_t>(nr, 1) * 4);
    require(out.frag_off && out.dataset, MCZI_E_NOMEM, "synthetic reads: out of memory");
    out.frag_off[0] = 0;
    for (size_t i = 0; i < nr; i++) {
        out.frag_off[i + 1] = out.frag_off[i] + rlen[i];
        out.dataset[i] = reads[i].dataset;
    }
    const uint64_t total = out.frag_off[nr];
    out.n_bases = total;
    out.n_words = (total + 15) / 16 + 32;
    out.words = (uint32_t *)calloc(out.n_words, 4);
    require(out.words != nullptr, MCZI_E_NOMEM, "synthetic reads: out of memory");
    // pack in parallel: each thread's bases are one contiguous range; only the
    // words at its two ends can be shared with a neighbour (atomic OR there)
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; t++) {
            th.emplace_back([&, t]() {
                size_t b, e;
                rng_of(t, b, e);
                const uint64_t p0 = out.frag_off[b], p1 = out.frag_off[e];
                if (p0 == p1) return;
                const uint64_t wf = p0 >> 4, wl = (p1 - 1) >> 4;
                const std::vector<uint8_t> &src = tbuf[t];
                uint32_t acc = 0;
                uint64_t w = wf;
                for (uint64_t p = p0; p < p1; p++) {
                    if ((p >> 4) != w) {
                        if (w == wf || w == wl) __atomic_fetch_or(&out.words[w], acc, __ATOMIC_RELAXED);
                        else out.words[w] = acc;
                        w = p >> 4;
                        acc = 0;
                    }
                    acc |= (uint32_t)src[p - p0] << ((p & 15u) * 2u);
                }
                __atomic_fetch_or(&out.words[w], acc, __ATOMIC_RELAXED);
                std::vector<uint8_t>().swap(tbuf[t]);
            });
        }
        for (auto &x : th) x.join();
    }
}

void synth_reads_from(const uint32_t *const *gw, const uint64_t *gl, uint32_t ng, uint64_t n_reads,
                      uint32_t read_len, double sub, uint64_t seed, mczi_synth_reads &out) {
    require(ng > 0, MCZI_E_INVALID, "at least one genome is required");
    std::vector<std::vector<uint8_t>> genomes(ng);
    for (uint32_t g = 0; g < ng; g++) {
        genomes[g].resize(gl[g]);
        for (uint64_t p = 0; p < gl[g]; p++) genomes[g][p] = (uint8_t)((gw[g][p >> 4] >> ((p & 15u) * 2u)) & 3u);
        require(gl[g] >= read_len, MCZI_E_INVALID, "genome shorter than the read length");
    }
    const uint64_t total = n_reads * read_len;
    out.n_bases = total;
    out.n_frag = n_reads;
    out.n_words = (total + 15) / 16 + 32;
    out.words = (uint32_t *)calloc(out.n_words, 4);
    out.frag_off = (uint64_t *)malloc((n_reads + 1) * 8);
    out.dataset = (uint32_t *)malloc(std::max<uint64_t>(n_reads, 1) * 4);
    require(out.words && out.frag_off && out.dataset, MCZI_E_NOMEM, "synthetic reads: out of memory");
    for (uint64_t i = 0; i <= n_reads; i++) out.frag_off[i] = i * read_len;
    unsigned nt = host_workers(size_t(1) << 20);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([&, t]() {
            std::vector<uint8_t> tmp;
            // threads own disjoint word ranges: read_len * reads-per-thread aligned to 16 bases
            const uint64_t per = ((n_reads + nt - 1) / nt + 15) / 16 * 16;
            const uint64_t r0 = std::min<uint64_t>(n_reads, per * t), r1 = std::min<uint64_t>(n_reads, r0 + per);
            for (uint64_t i = r0; i < r1; i++) {
                SplitMix64 rng(read_seed(seed, i));
                sample_read(genomes[i % ng], read_len, sub, 0.0, 0.0, rng, tmp);
                uint64_t p = i * read_len;
                for (uint8_t c : tmp) {
                    out.words[p >> 4] |= (uint32_t)c << ((p & 15u) * 2u);
                    p++;
                }
                out.dataset[i] = (uint32_t)(i % ng);
            }
        });
    }
    for (auto &x : th) x.join();
}

void synth_write_fasta(const mczi_synth_reads &r, const std::string &path, int64_t ds_filter) {
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) throw Error(MCZI_E_IO, "failed to create " + path);
    static const char kAscii[4] = {'A', 'C', 'T', 'G'};  // packed-seq code -> base
    std::string buf;
    buf.reserve(1 << 22);
    for (uint64_t i = 0; i < r.n_frag; i++) {
        if (ds_filter >= 0 && r.dataset && r.dataset[i] != (uint32_t)ds_filter) continue;
        buf += ">r" + std::to_string(i) + "\n";
        for (uint64_t p = r.frag_off[i]; p < r.frag_off[i + 1]; p++)
            buf.push_back(kAscii[(r.words[p >> 4] >> ((p & 15u) * 2u)) & 3u]);
        buf.push_back('\n');
        if (buf.size() >= (1u << 22)) {
            fwrite(buf.data(), 1, buf.size(), fp);
            buf.clear();
        }
    }
    fwrite(buf.data(), 1, buf.size(), fp);
    if (fclose(fp) != 0) throw Error(MCZI_E_IO, "failed to write " + path);
}

void synth_genome_words(uint64_t len, uint64_t seed, uint32_t **words) {
    std::vector<uint8_t> g = make_genome(len, seed);
    uint32_t *w = (uint32_t *)calloc((len + 15) / 16 + 32, 4);
    require(w != nullptr, MCZI_E_NOMEM, "synthetic genome: out of memory");
    for (uint64_t p = 0; p < len; p++) w[p >> 4] |= (uint32_t)g[p] << ((p & 15u) * 2u);
    *words = w;
}

}  // namespace mczi
