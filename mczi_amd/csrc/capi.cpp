// C ABI (include/mczi.h): argument checks, error capture, and the drivers
// that mirror the reference's lib.rs / zi.rs entry points.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <sys/resource.h>

#include "internal.hpp"
#include "zi.hpp"

using namespace mczi;

namespace {

thread_local std::string g_err;

template <typename F>
int guarded(const char *ctx, F &&f) {
    try {
        f();
        g_err.clear();
        return MCZI_OK;
    } catch (const Error &e) {
        g_err = std::string(ctx) + ": " + e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = std::string(ctx) + ": out of host memory";
        return MCZI_E_NOMEM;
    } catch (const std::exception &e) {
        g_err = std::string(ctx) + ": " + e.what();
        return MCZI_E_INVALID;
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// validate_shape_config / validate_config / validate_dataset_presence_config
// (src/lib.rs:110-152), same messages.
void validate_shape(const mczi_counter_config &c) {
    require(c.k > 0, MCZI_E_INVALID, "k must be greater than 0");
    require(c.k <= 64, MCZI_E_INVALID, "k must be <= 64 because MC stores k-mers in u128");
    require(c.minimizer > 0, MCZI_E_INVALID, "minimizer size must be greater than 0");
    require(c.minimizer <= c.k, MCZI_E_INVALID, "minimizer size must be <= k");
    require(c.minimizer <= 64, MCZI_E_INVALID,
            "minimizer size must be <= 64 because MC stores minimizers in u128");
    require(c.partition_count > 0, MCZI_E_INVALID, "partition count must be greater than 0");
}
void validate_normal(const mczi_counter_config &c) {
    validate_shape(c);
    require(c.threshold < 255, MCZI_E_INVALID,
            "threshold must be < 255 because compact minimizer counts saturate in u8");
}
void validate_datasets(const mczi_counter_config &c) {
    validate_shape(c);
    require(c.k <= 32, MCZI_E_INVALID, "FOFN dataset-presence mode currently requires k <= 32");
    require(c.threshold < 0x7fffffffull, MCZI_E_INVALID, "FOFN dataset-presence threshold must be < 2^31");
}

std::vector<std::string> to_paths(const char *const *paths, size_t n) {
    std::vector<std::string> v;
    for (size_t i = 0; i < n; i++) {
        require(paths && paths[i], MCZI_E_INVALID, "null input path");
        v.emplace_back(paths[i]);
    }
    return v;
}

bool is_datasets(int mode) { return mode == MCZI_MODE_DATASETS || mode == MCZI_MODE_DATASETS_TWO_PASS; }

std::unique_ptr<Batch> load_inputs(const std::vector<std::string> &paths, uint32_t k, bool datasets) {
    HostReads r;
    r.n_datasets = datasets ? (uint32_t)paths.size() : 1;
    for (size_t i = 0; i < paths.size(); i++) parse_fastx_file(paths[i], k, datasets ? (uint32_t)i : 0, r);
    return make_batch(std::move(r));
}

std::vector<mczi_counted_kmer> run_count(const std::vector<std::string> &paths,
                                         const mczi_counter_config &cfg, int mode, int order,
                                         mczi_run_stats *st) {
    auto b = load_inputs(paths, cfg.k, is_datasets(mode));
    std::vector<mczi_counted_kmer> out;
    if (b->n_frag == 0) {
        if (st) *st = mczi_run_stats{};
        return out;
    }
    batch_count(*b, cfg, mode, MCZI_LAUNCH_BATCHED, order, st);
    batch_download(*b, out);
    return out;
}

void copy_out(const std::vector<mczi_counted_kmer> &v, mczi_counted_kmer **out, size_t *n_out) {
    require(out && n_out, MCZI_E_INVALID, "null output pointer");
    *n_out = v.size();
    *out = (mczi_counted_kmer *)malloc(std::max<size_t>(v.size(), 1) * sizeof(mczi_counted_kmer));
    require(*out != nullptr, MCZI_E_NOMEM, "out of host memory");
    if (!v.empty()) memcpy(*out, v.data(), v.size() * sizeof(mczi_counted_kmer));
}

// The count paths stage every super-k-mer (single pass): the reference's
// pass-1 filter cannot change the k-mers or counts emitted. Calls that report
// the minimizer statistics run pass 1 on the device (two_pass; *st gets the
// run's statistics).
std::vector<mczi_counted_kmer> count_inputs_impl(const char *const *paths, size_t n,
                                                 const mczi_counter_config *cfg, int order,
                                                 bool two_pass = false, mczi_run_stats *st = nullptr) {
    require(cfg != nullptr, MCZI_E_INVALID, "null config");
    validate_normal(*cfg);
    require(n > 0, MCZI_E_INVALID, "at least one input file is required");
    const double t0 = now_s();
    mczi_run_stats local{};
    auto v = run_count(to_paths(paths, n), *cfg, two_pass ? MCZI_MODE_TWO_PASS : MCZI_MODE_SINGLE_PASS, order,
                       st ? st : &local);
    log_phase("3_kmer_counting_and_output", now_s() - t0);
    return v;
}

std::vector<mczi_counted_kmer> count_datasets_impl(const char *const *paths, size_t n,
                                                   const mczi_counter_config *cfg, bool two_pass = false,
                                                   mczi_run_stats *st = nullptr) {
    require(cfg != nullptr, MCZI_E_INVALID, "null config");
    validate_datasets(*cfg);
    require(n > 0, MCZI_E_INVALID, "at least one dataset file is required");
    if (st) *st = mczi_run_stats{};
    if (cfg->threshold >= n) {  // src/lib.rs:644-648
        log_phase("1_dataset_minimizer_counting", 0);
        log_phase("2_dataset_kmer_counting", 0);
        return {};
    }
    const double t0 = now_s();
    auto v = run_count(to_paths(paths, n), *cfg, two_pass ? MCZI_MODE_DATASETS_TWO_PASS : MCZI_MODE_DATASETS,
                       MCZI_ORDER_ENCODED, st);
    log_phase("2_dataset_kmer_counting", now_s() - t0);
    return v;
}

// ---- MCZI in one process (src/mczi.rs:238-451) ------------------------------
// log_resource_phase (src/lib.rs:1012-1021): "MCZI phaseN detail W.WWWs CPU
// C.CCCs NMB RAM", CPU = user + system time of the phase, RAM = peak RSS.
class ResourcePhase {
  public:
    ResourcePhase() : t0_(now_s()), cpu0_(cpu_s()) {}
    void log(const char *name) const {
        rusage u{};
        getrusage(RUSAGE_SELF, &u);
        const unsigned long long mb = ((unsigned long long)u.ru_maxrss + 1023) / 1024;
        fprintf(stderr, "MCZI %s %.3fs CPU %.3fs %lluMB RAM\n", phase_display_name(name).c_str(), now_s() - t0_,
                cpu_s() - cpu0_, mb);
    }
    static void zero(const char *name) {
        rusage u{};
        getrusage(RUSAGE_SELF, &u);
        fprintf(stderr, "MCZI %s 0.000s CPU 0.000s %lluMB RAM\n", phase_display_name(name).c_str(),
                ((unsigned long long)u.ru_maxrss + 1023) / 1024);
    }

  private:
    static double cpu_s() {
        rusage u{};
        getrusage(RUSAGE_SELF, &u);
        return u.ru_utime.tv_sec + u.ru_stime.tv_sec + 1e-6 * (u.ru_utime.tv_usec + u.ru_stime.tv_usec);
    }
    double t0_, cpu0_;
};

// log_mczi_stat with format_stat_u64's comma groups (src/mczi.rs:1600-1632)
void mczi_stat(const char *name, uint64_t v) {
    const std::string d = std::to_string(v);
    std::string g;
    for (size_t i = 0; i < d.size(); i++) {
        if (i && (d.size() - i) % 3 == 0) g.push_back(',');
        g.push_back(d[i]);
    }
    fprintf(stderr, "MCZI_STAT\t%s\t%s\n", name, g.c_str());
}

// The index inputs are counted by MC and the counted k-mers become the
// device membership index directly (the reference goes through GGCAT
// simplitigs and SSHash: same membership); the queries are subtracted as in
// ZI. Returns the query scan with the sorted unique absent k-mers.
// run()'s argument checks (src/mczi.rs:239-311), before any device work.
void mczi_check(const char *const *ip, size_t ni, bool datasets, const char *const *qp, size_t nq,
                const mczi_counter_config &cfg) {
    require(cfg.k >= 3 && cfg.k <= 63, MCZI_E_INVALID, "k must be in [3, 63] for sshash-rs");
    require(cfg.k % 2 == 1, MCZI_E_INVALID, "k must be odd for sshash-rs");
    require(cfg.minimizer > 0 && cfg.minimizer < cfg.k, MCZI_E_INVALID, "minimizer size must be > 0 and < k");
    require(cfg.partition_count > 0, MCZI_E_INVALID, "--partition-count must be greater than 0");
    require(ni > 0 && ip != nullptr, MCZI_E_INVALID, "at least one index input is required");
    require(nq > 0 && qp != nullptr, MCZI_E_INVALID, "at least one query input is required");
    if (datasets)
        validate_datasets(cfg);
    else
        validate_normal(cfg);
}

ZiScan mczi_run(const char *const *ip, size_t ni, bool datasets, const char *const *qp, size_t nq,
                const mczi_counter_config &cfg, bool no_output, mczi_mczi_stats *stats) {
    Device &d = Device::get();
    std::unique_ptr<Batch> b;
    uint64_t n_index = 0;
    if (datasets && cfg.threshold >= ni) {  // src/mczi.rs:325-337
        fprintf(stderr,
                "MCZI_WARN\tindex_fofn_threshold_excludes_all_kmers\tindex_inputs\t%zu\tthreshold\t%llu\tmessage\t"
                "no k-mer can pass a dataset-presence threshold greater than or equal to the number of index "
                "datasets\n",
                ni, (unsigned long long)cfg.threshold);
        ResourcePhase::zero("1_index_minimizer_presence_counting");
        ResourcePhase::zero("2_index_kmer_partition_counting");
        mczi_stat("index_minimizers_above_threshold", 0);
    } else {
        ResourcePhase ph;
        b = load_inputs(to_paths(ip, ni), cfg.k, datasets);
        mczi_run_stats st{};
        if (b->n_frag)
            batch_count(*b, cfg, datasets ? MCZI_MODE_DATASETS_TWO_PASS : MCZI_MODE_TWO_PASS, MCZI_LAUNCH_BATCHED,
                        MCZI_ORDER_ENCODED, &st);
        n_index = b->n_frag ? b->n_out : 0;
        // KmerFastaWriteStats.minimizers_above_threshold of the index count
        // (src/mczi.rs:351-354)
        mczi_stat("index_minimizers_above_threshold", datasets ? st.minimizers_retained : st.unique_minimizers);
        if (datasets) {
            ResourcePhase::zero("1_index_minimizer_presence_counting");
            ph.log("2_index_kmer_partition_counting");
        } else {
            ph.log("1_index_kmer_counting");
        }
    }
    mczi_stat("index_kmers_above_threshold", n_index);
    ResourcePhase::zero("3_ggcat_simplitigs");  // no simplitig round trip: the k-mers go straight in
    ZiIndex idx;
    {
        ResourcePhase ph;
        zi_build_index_from_keys(idx, n_index ? b->out_kmer.as<unsigned long long>() : nullptr,
                                 n_index && cfg.k > 32 ? b->out_hi.as<unsigned long long>() : nullptr, n_index,
                                 cfg.k, cfg.minimizer, d.stream());
        ph.log("4_sshash_indexing");
    }
    b.reset();
    ResourcePhase ph;
    ZiScan scan = zi_subtract_files(idx, qp, nq, d.stream());
    ph.log(no_output ? "5_query_subtraction_no_output" : "5_query_subtraction");
    // log_query_stats (src/mczi.rs:1634-1650)
    mczi_stat("query_kmers_scanned", scan.windows);
    mczi_stat("query_kmers_filtered_by_zi", scan.windows - scan.absent_windows);
    mczi_stat("query_kmers_not_filtered_by_zi", scan.absent_windows);
    if (stats) {
        stats->index_kmers = n_index;
        stats->query_kmers_scanned = scan.windows;
        stats->query_kmers_filtered = scan.windows - scan.absent_windows;
        stats->query_kmers_not_filtered = scan.absent_windows;
        stats->query_unique_kmers_not_filtered = scan.absent.n;
    }
    return scan;
}

}  // namespace

extern "C" {

int mczi_mczi_subtract(const char *const *index_paths, size_t n_index, int index_datasets,
                       const char *const *query_paths, size_t n_query, const mczi_counter_config *cfg,
                       const char *out_path, mczi_mczi_stats *stats) {
    return guarded("MCZI failed", [&] {
        require(cfg != nullptr, MCZI_E_INVALID, "null config");
        mczi_check(index_paths, n_index, index_datasets != 0, query_paths, n_query, *cfg);
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        ResourcePhase total;
        const bool no_output = out_path == nullptr;
        ZiScan scan = mczi_run(index_paths, n_index, index_datasets != 0, query_paths, n_query, *cfg, no_output, stats);
        if (no_output) {
            ResourcePhase::zero("6_no_output");
        } else {
            ResourcePhase ph;
            ZiKeys &a = scan.absent;
            write_simplitig_fasta_keys(out_path, a.n, cfg->k, [&](uint64_t *lo, uint64_t *hi) {
                if (a.n) memcpy(lo, a.lo, a.n * 8);
                if (hi) {
                    if (a.hi && a.n)
                        memcpy(hi, a.hi, a.n * 8);
                    else if (a.n)
                        memset(hi, 0, a.n * 8);
                }
            });
            ph.log("6_ggcat_simplitig_output");
        }
        total.log("total");
    });
}

int mczi_mczi_absent(const char *const *index_paths, size_t n_index, int index_datasets,
                     const char *const *query_paths, size_t n_query, const mczi_counter_config *cfg,
                     mczi_counted_kmer **out, size_t *n_out, mczi_mczi_stats *stats) {
    return guarded("MCZI failed", [&] {
        require(cfg != nullptr && out != nullptr && n_out != nullptr, MCZI_E_INVALID, "null argument");
        mczi_check(index_paths, n_index, index_datasets != 0, query_paths, n_query, *cfg);
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        ZiScan scan = mczi_run(index_paths, n_index, index_datasets != 0, query_paths, n_query, *cfg, true, stats);
        copy_out(scan.absent.counted(), out, n_out);
    });
}

const char *mczi_last_error(void) { return g_err.c_str(); }

int mczi_set_threads(uint32_t n) {
    g_host_threads.store(n, std::memory_order_relaxed);
    return MCZI_OK;
}
const char *mczi_version(void) { return "mczi-hip 0.1.0 (gfx950)"; }
void mczi_free(void *p) { free(p); }

static void reads_out(const HostReads &r, uint32_t **words, uint64_t **frag_off, size_t *n_frag) {
        const size_t nw = (size_t)((r.n_bases() + 15) / 16);
        *words = (uint32_t *)malloc(std::max<size_t>(nw, 1) * 4);
        *frag_off = (uint64_t *)malloc(r.off.size() * 8);
        if (!*words || !*frag_off) {
            free(*words);
            free(*frag_off);
            *words = nullptr;
            *frag_off = nullptr;
            throw Error(MCZI_E_NOMEM, "out of host memory");
        }
        if (nw) memcpy(*words, r.words.data(), nw * 4);
        memcpy(*frag_off, r.off.data(), r.off.size() * 8);
        *n_frag = (size_t)r.n_frag();
}

int mczi_pack_fastx(const char *path, uint32_t min_len, uint32_t **words, uint64_t **frag_off,
                    size_t *n_frag) {
    return guarded("failed to read FASTA/FASTQ", [&] {
        require(path && words && frag_off && n_frag, MCZI_E_INVALID, "null argument");
        HostReads r;
        parse_fastx_file(path, min_len, 0, r);
        reads_out(r, words, frag_off, n_frag);
    });
}

int mczi_write_read_cache(const char *fastx_path, uint32_t min_len, const char *cache_path) {
    return guarded("failed to write packed-read cache", [&] {
        require(fastx_path && cache_path, MCZI_E_INVALID, "null argument");
        HostReads r;
        parse_fastx_file(fastx_path, min_len, 0, r);
        write_read_cache(cache_path, r);
    });
}

int mczi_read_read_cache(const char *cache_path, uint32_t min_len, uint32_t **words, uint64_t **frag_off,
                         size_t *n_frag) {
    return guarded("failed to read packed-read cache", [&] {
        require(cache_path && words && frag_off && n_frag, MCZI_E_INVALID, "null argument");
        HostReads r;
        read_read_cache(cache_path, min_len, r);
        reads_out(r, words, frag_off, n_frag);
    });
}

int mczi_read_superkmer_partitions(const char *const *paths, size_t n, uint32_t min_len, uint32_t **words,
                                   uint64_t **frag_off, size_t *n_frag) {
    return guarded("failed to read super-k-mer partitions", [&] {
        require((n == 0 || paths) && words && frag_off && n_frag, MCZI_E_INVALID, "null argument");
        HostReads r;
        for (size_t i = 0; i < n; i++) read_superkmer_file(paths[i], min_len, r);
        reads_out(r, words, frag_off, n_frag);
    });
}

int mczi_write_superkmer_partitions(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                    const char *dir) {
    return guarded("failed to write super-k-mer partitions", [&] {
        require(cfg && dir && (n == 0 || paths), MCZI_E_INVALID, "null argument");
        validate_normal(*cfg);
        require(n > 0, MCZI_E_INVALID, "at least one input file is required");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        auto b = load_inputs(to_paths(paths, n), cfg->k, false);
        std::vector<uint64_t> rec;
        std::vector<uint32_t> part;
        uint32_t rec_words = cfg->k > 32 ? 4 : 2, p_int = 0;
        if (b->n_frag) {
            batch_count(*b, *cfg, MCZI_MODE_TWO_PASS, MCZI_LAUNCH_BATCHED, MCZI_ORDER_NONE, nullptr);
            batch_download_records(*b, rec, part, rec_words, p_int);
        }
        write_superkmer_partitions(dir, cfg->partition_count, rec, part, rec_words);
    });
}

int mczi_count_superkmer_partitions(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                    mczi_counted_kmer **out, size_t *n_out) {
    return guarded("failed to count super-k-mer partitions", [&] {
        require(cfg && out && n_out && (n == 0 || paths), MCZI_E_INVALID, "null argument");
        validate_normal(*cfg);
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        HostReads r;
        for (size_t i = 0; i < n; i++) read_superkmer_file(paths[i], cfg->k, r);
        auto b = make_batch(std::move(r));
        std::vector<mczi_counted_kmer> v;
        if (b->n_frag) {
            batch_count(*b, *cfg, MCZI_MODE_TWO_PASS, MCZI_LAUNCH_BATCHED, MCZI_ORDER_ENCODED, nullptr);
            batch_download(*b, v);
        }
        copy_out(v, out, n_out);
    });
}

int mczi_write_counted(const char *path, const mczi_counted_kmer *kmers, size_t n) {
    return guarded("failed to write counted k-mers", [&] {
        require(path && (kmers || n == 0), MCZI_E_INVALID, "null argument");
        write_counted_file(path, kmers, n);
    });
}

int mczi_read_counted(const char *path, mczi_counted_kmer **out, size_t *n_out) {
    return guarded("failed to read counted k-mers", [&] {
        require(path && out && n_out, MCZI_E_INVALID, "null argument");
        copy_out(read_counted_file(path), out, n_out);
    });
}

int mczi_validate_config(const mczi_counter_config *cfg, int datasets) {
    return guarded("invalid configuration", [&] {
        require(cfg != nullptr, MCZI_E_INVALID, "null config");
        if (datasets) validate_datasets(*cfg);
        else validate_normal(*cfg);
    });
}

// expand_fofns (src/lib.rs:242-276)
int mczi_expand_fofns(const char *const *fofns, size_t n, char ***out, size_t *n_out) {
    return guarded("failed to expand FOFN", [&] {
        require(out && n_out, MCZI_E_INVALID, "null output pointer");
        std::vector<std::string> inputs;
        for (size_t i = 0; i < n; i++) {
            std::string f(fofns[i]);
            std::ifstream in(f);
            require((bool)in, MCZI_E_IO, "failed to read FOFN " + f);
            std::string base = ".";
            size_t sl = f.rfind('/');
            if (sl != std::string::npos) base = sl == 0 ? "/" : f.substr(0, sl);
            std::string line;
            while (std::getline(in, line)) {
                size_t a = line.find_first_not_of(" \t\r\n");
                if (a == std::string::npos) continue;
                size_t b = line.find_last_not_of(" \t\r\n");
                std::string t = line.substr(a, b - a + 1);
                if (t.empty() || t[0] == '#') continue;
                if (t[0] != '/') t = (base == "/" ? "/" : base + "/") + t;
                inputs.push_back(t);
            }
        }
        require(!inputs.empty(), MCZI_E_INVALID, "no inputs found in FOFN(s)");
        char **arr = (char **)calloc(inputs.size(), sizeof(char *));
        require(arr != nullptr, MCZI_E_NOMEM, "out of host memory");
        for (size_t i = 0; i < inputs.size(); i++) arr[i] = strdup(inputs[i].c_str());
        *out = arr;
        *n_out = inputs.size();
    });
}

void mczi_free_strings(char **strs, size_t n) {
    if (!strs) return;
    for (size_t i = 0; i < n; i++) free(strs[i]);
    free(strs);
}

int mczi_count_inputs(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                      mczi_counted_kmer **out, size_t *n_out) {
    return guarded("failed to count inputs", [&] {
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        copy_out(count_inputs_impl(paths, n, cfg, MCZI_ORDER_ENCODED), out, n_out);
    });
}

int mczi_count_datasets(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                        mczi_counted_kmer **out, size_t *n_out) {
    return guarded("failed to count datasets", [&] {
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        copy_out(count_datasets_impl(paths, n, cfg), out, n_out);
    });
}

// Count, then hand only the sorted k-mer words to the simplitig writer (the
// counts are not written to simplitig FASTA).
void count_to_simplitigs(const std::vector<std::string> &paths, const mczi_counter_config &cfg, int mode,
                         const char *out_path, const char *count_phase) {
    const double t0 = now_s();
    auto b = load_inputs(paths, cfg.k, is_datasets(mode));
    if (b->n_frag) batch_count(*b, cfg, mode, MCZI_LAUNCH_BATCHED, MCZI_ORDER_ENCODED, nullptr);
    log_phase(count_phase, now_s() - t0);
    const size_t n_out = b->n_frag ? (size_t)b->n_out : 0;
    const double t1 = now_s();
    write_simplitig_fasta_keys(
        out_path, n_out, cfg.k, [&](uint64_t *lo, uint64_t *hi) { batch_download_keys(*b, lo, hi); },
        [&](uint64_t *link) {
            device_forced_links(b->out_kmer.as<unsigned long long>(),
                                cfg.k > 32 ? b->out_hi.as<unsigned long long>() : nullptr, n_out, cfg.k, link);
        });
    log_phase(is_datasets(mode) ? "3_output_streaming" : "4_simplitig_output", now_s() - t1);
}

int mczi_count_inputs_to_fasta_path(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                    const char *out_path) {
    return guarded("failed to write simplitig FASTA", [&] {
        require(out_path != nullptr && cfg != nullptr, MCZI_E_INVALID, "null argument");
        validate_normal(*cfg);
        require(n > 0, MCZI_E_INVALID, "at least one input file is required");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        const double t0 = now_s();
        count_to_simplitigs(to_paths(paths, n), *cfg, MCZI_MODE_SINGLE_PASS, out_path, "3_kmer_counting_and_output");
        log_phase("fasta_total", now_s() - t0);
    });
}

int mczi_count_datasets_to_fasta_path(const char *const *paths, size_t n,
                                      const mczi_counter_config *cfg, const char *out_path) {
    return guarded("failed to write simplitig FASTA", [&] {
        require(out_path != nullptr && cfg != nullptr, MCZI_E_INVALID, "null argument");
        validate_datasets(*cfg);
        require(n > 0, MCZI_E_INVALID, "at least one dataset file is required");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        const double t0 = now_s();
        if (cfg->threshold >= n) {  // src/lib.rs:644-648
            log_phase("1_dataset_minimizer_counting", 0);
            log_phase("2_dataset_kmer_counting", 0);
            const double t1 = now_s();
            write_simplitig_fasta_keys(out_path, 0, cfg->k, [](uint64_t *, uint64_t *) {});
            log_phase("3_output_streaming", now_s() - t1);
        } else {
            count_to_simplitigs(to_paths(paths, n), *cfg, MCZI_MODE_DATASETS, out_path, "2_dataset_kmer_counting");
        }
        log_phase("total", now_s() - t0);
    });
}

int mczi_count_inputs_to_kff_path(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                  const char *out_path) {
    return guarded("failed to write KFF", [&] {
        require(cfg != nullptr && out_path != nullptr, MCZI_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        validate_normal(*cfg);
        require(n > 0, MCZI_E_INVALID, "at least one input file is required");
        const double t0 = now_s();
        if (cfg->k <= 32) {
            // KMC-style single pass, u8 counts saturating at min(x+1, 255)
            // (src/lib.rs:599-635, 2489-2506)
            mczi_run_stats st{};
            auto v = run_count(to_paths(paths, n), *cfg, MCZI_MODE_SINGLE_PASS, MCZI_ORDER_KFF_U8, &st);
            fprintf(stderr, "MC_STAT\tpartition_bytes\t%llu\n", (unsigned long long)st.partition_bytes);
            log_phase("1_superkmer_partitioning", st.ms_scan1 / 1e3);
            const double t1 = now_s();
            write_kff(out_path, v, *cfg, 1, false, std::min<uint64_t>(cfg->threshold + 1, 255));
            log_phase("2b_output_streaming", now_s() - t1);
        } else {
            auto v = count_inputs_impl(paths, n, cfg, MCZI_ORDER_ENCODED);
            write_kff(out_path, v, *cfg, 8, false, 0);
        }
        log_phase("total", now_s() - t0);
    });
}

int mczi_count_datasets_to_kff_path(const char *const *paths, size_t n,
                                    const mczi_counter_config *cfg, const char *out_path) {
    return guarded("failed to write KFF", [&] {
        require(out_path != nullptr, MCZI_E_INVALID, "null output path");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        const double t0 = now_s();
        auto v = count_datasets_impl(paths, n, cfg);
        const double t1 = now_s();
        write_kff(out_path, v, *cfg, 4, false, 0);
        log_phase("3_output_streaming", now_s() - t1);
        log_phase("total", now_s() - t0);
    });
}

int mczi_count_inputs_to_kmer_fasta_path(const char *const *paths, size_t n,
                                         const mczi_counter_config *cfg, const char *out_path,
                                         mczi_kmer_fasta_stats *stats) {
    return guarded("failed to write k-mer FASTA", [&] {
        require(out_path != nullptr, MCZI_E_INVALID, "null output path");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        const double t0 = now_s();
        mczi_run_stats st{};
        auto v = count_inputs_impl(paths, n, cfg, MCZI_ORDER_ENCODED, stats != nullptr, &st);
        write_kmer_fasta(out_path, v, cfg->k);
        if (stats) {
            // retained_minimizers = unique_hashes() of pass 1 (src/lib.rs:950-951, 326-351)
            stats->selected_kmers = v.size();
            stats->minimizers_above_threshold = st.unique_minimizers;
            stats->total_seconds = now_s() - t0;
        }
    });
}

int mczi_count_datasets_to_kmer_fasta_path(const char *const *paths, size_t n,
                                           const mczi_counter_config *cfg, const char *out_path,
                                           mczi_kmer_fasta_stats *stats) {
    return guarded("failed to write k-mer FASTA", [&] {
        require(out_path != nullptr, MCZI_E_INVALID, "null output path");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        const double t0 = now_s();
        mczi_run_stats st{};
        auto v = count_datasets_impl(paths, n, cfg, stats != nullptr, &st);
        write_kmer_fasta(out_path, v, cfg->k);
        if (stats) {
            // the minimizers retained after the last dataset (src/lib.rs:759-760, 815)
            stats->selected_kmers = v.size();
            stats->minimizers_above_threshold = st.minimizers_retained;
            stats->total_seconds = now_s() - t0;
        }
    });
}

int mczi_run_inputs_phase12_direct(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                   mczi_phase12_stats *stats) {
    return guarded("failed to run phases 1-2", [&] {
        require(cfg != nullptr && stats != nullptr, MCZI_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        validate_normal(*cfg);
        require(n > 0, MCZI_E_INVALID, "at least one input file is required");
        const double t0 = now_s();
        auto b = load_inputs(to_paths(paths, n), cfg->k, false);
        mczi_run_stats st{};
        if (b->n_frag) batch_count(*b, *cfg, MCZI_MODE_TWO_PASS, MCZI_LAUNCH_BATCHED, MCZI_ORDER_NONE, &st);
        *stats = mczi_phase12_stats{};
        stats->phase1 = st.ms_pass1 / 1e3;
        stats->phase2 = (st.ms_total > 0 ? st.ms_total : 0) / 1e3;
        stats->unique_minimizer_hashes = st.unique_minimizers;
        stats->partition_bytes = st.partition_bytes;
        log_phase("1_minimizer_counting", stats->phase1);
        log_phase("2_superkmer_partitioning", stats->phase2);
        stats->total = now_s() - t0;
        log_phase("phase12_total", stats->total);
    });
}

int mczi_count_inputs_minimizer_phase_direct(const char *const *paths, size_t n, const mczi_counter_config *cfg,
                                             uint64_t *unique) {
    return guarded("failed to count minimizers", [&] {
        require(cfg != nullptr && unique != nullptr, MCZI_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        validate_normal(*cfg);
        require(n > 0, MCZI_E_INVALID, "at least one input file is required");
        auto b = load_inputs(to_paths(paths, n), cfg->k, false);
        *unique = b->n_frag ? batch_minimizer_count(*b, *cfg) : 0;
    });
}

int mczi_write_fasta_path(const char *path, const mczi_counted_kmer *kmers, size_t n, uint32_t k) {
    return guarded("failed to write FASTA", [&] {
        require(kmers != nullptr || n == 0, MCZI_E_INVALID, "null k-mer array");
        write_simplitig_fasta(path, kmers, n, k, 0, nullptr);
    });
}

int mczi_write_fasta_after_index(const char *path, const mczi_counted_kmer *kmers, size_t n, uint32_t k,
                                 size_t last_written_index, size_t *last_index) {
    return guarded("failed to write FASTA", [&] {
        require(kmers != nullptr || n == 0, MCZI_E_INVALID, "null k-mer array");
        write_simplitig_fasta(path, kmers, n, k, last_written_index, last_index);
    });
}

static void strings_out(const std::vector<std::string> &v, char ***seqs, size_t *n_seqs) {
    char **arr = (char **)calloc(std::max<size_t>(v.size(), 1), sizeof(char *));
    require(arr != nullptr, MCZI_E_NOMEM, "out of host memory");
    for (size_t i = 0; i < v.size(); i++) {
        arr[i] = strdup(v[i].c_str());
        if (!arr[i]) {
            mczi_free_strings(arr, i);
            throw Error(MCZI_E_NOMEM, "out of host memory");
        }
    }
    *seqs = arr;
    *n_seqs = v.size();
}

int mczi_simplitig_sequences(const mczi_counted_kmer *kmers, size_t n, uint32_t k, char ***seqs,
                             size_t *n_seqs) {
    return guarded("failed to build simplitigs", [&] {
        require((kmers != nullptr || n == 0) && seqs && n_seqs, MCZI_E_INVALID, "null argument");
        strings_out(simplitig_sequences(kmers, n, k), seqs, n_seqs);
    });
}

int mczi_simplitig_sequences_parallel(const mczi_counted_kmer *kmers, size_t n, uint32_t k,
                                      uint32_t minimizer, uint32_t partition_count, char ***seqs,
                                      size_t *n_seqs) {
    return guarded("failed to build simplitigs", [&] {
        require((kmers != nullptr || n == 0) && seqs && n_seqs, MCZI_E_INVALID, "null argument");
        strings_out(simplitig_sequences_parallel(kmers, n, k, minimizer, partition_count), seqs, n_seqs);
    });
}

int mczi_write_kff_path(const char *path, const mczi_counted_kmer *kmers, size_t n,
                        const mczi_counter_config *cfg) {
    return guarded("failed to write KFF", [&] {
        require(cfg != nullptr, MCZI_E_INVALID, "null config");
        validate_normal(*cfg);
        std::vector<mczi_counted_kmer> v(kmers, kmers + n);
        write_kff(path, v, *cfg, 8, true, 0);
    });
}

int mczi_decode_kmer(uint64_t lo, uint64_t hi, uint32_t k, char *out) {
    return guarded("failed to decode k-mer", [&] {
        require(k > 0 && k <= 64 && out != nullptr, MCZI_E_INVALID, "k must be in 1..=64");
        static const char a[4] = {'A', 'C', 'G', 'T'};
        for (uint32_t i = 0; i < k; i++) {
            const uint32_t sh = 2 * (k - 1 - i);
            const uint64_t v = sh >= 64 ? (hi >> (sh - 64)) : ((lo >> sh) | (sh ? (hi << (64 - sh)) : 0));
            out[i] = a[v & 3u];
        }
    });
}

int mczi_minimizer_run_hashes(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m,
                              uint32_t *hashes, uint32_t *starts, size_t *n_runs) {
    return guarded("failed to compute minimizer runs", [&] {
        require(n_runs != nullptr, MCZI_E_INVALID, "null output pointer");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        std::vector<uint32_t> h, s;
        device_run_hashes(packed, n_bases, k, m, h, s);
        if (!h.empty()) {
            require(hashes && starts, MCZI_E_INVALID, "null output arrays");
            memcpy(hashes, h.data(), h.size() * 4);
            memcpy(starts, s.data(), s.size() * 4);
        }
        *n_runs = h.size();
    });
}

int mczi_minimizer_run_hashes_noalloc(const uint8_t *packed, uint64_t n_bases, uint32_t k, uint32_t m,
                                      uint32_t *hashes, uint32_t *starts, size_t *n_runs) {
    return guarded("failed to compute minimizer runs", [&] {
        require(n_runs != nullptr, MCZI_E_INVALID, "null output pointer");
        require(k >= 1 && k <= 64 && m >= 1 && m <= k && k - m + 1 <= 64, MCZI_E_INVALID,
                "invalid k / m for the minimizer scan");
        require(k % 2 == 1, MCZI_E_INVALID, "window length l = k must be odd");
        require(n_bases <= NOALLOC_RUN_HASHES_MAX_BASES, MCZI_E_INVALID,
                "the allocation-free minimizer scan takes at most 16777216 bases per call");
        *n_runs = 0;
        if (n_bases < k) return;
        require(packed && hashes && starts, MCZI_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        device_run_hashes_noalloc(packed, n_bases, k, m, hashes, starts, n_runs);
    });
}

int mczi_copy_probe(uint64_t bytes, int reps, double *ms_per_copy) {
    return guarded("copy probe failed", [&] {
        require(ms_per_copy != nullptr, MCZI_E_INVALID, "null output pointer");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        *ms_per_copy = device_copy_probe(bytes, reps);
    });
}

int mczi_device_info(char *buf, size_t len) {
    return guarded("device query failed", [&] {
        require(buf != nullptr && len > 0, MCZI_E_INVALID, "null buffer");
        Device &d = Device::get();
        hipDeviceProp_t p;
        MCZI_HIP(hipGetDeviceProperties(&p, d.device()));
        int optin = 0;
        (void)hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, d.device());
        snprintf(buf, len,
                 "{\"name\": \"%s\", \"arch\": \"%s\", \"cus\": %d, \"clock_khz\": %d, "
                 "\"mem_clock_khz\": %d, \"bus_width\": %d, \"total_mem\": %zu, \"lds_per_block\": %zu, "
                 "\"lds_per_cu\": %zu, \"lds_per_block_optin\": %d, \"l2_bytes\": %d}",
                 p.name, p.gcnArchName, p.multiProcessorCount, p.clockRate, p.memoryClockRate,
                 p.memoryBusWidth, (size_t)p.totalGlobalMem, (size_t)p.sharedMemPerBlock,
                 (size_t)p.maxSharedMemoryPerMultiProcessor, optin, p.l2CacheSize);
    });
}

// ---- batched device API -----------------------------------------------------
struct mczi_batch {
    std::unique_ptr<Batch> b;
};

int mczi_batch_from_files(const char *const *paths, size_t n, uint32_t min_len, int one_dataset_per_file,
                          mczi_batch **out) {
    return guarded("failed to load batch", [&] {
        require(out != nullptr, MCZI_E_INVALID, "null output pointer");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        auto h = new mczi_batch();
        try {
            h->b = load_inputs(to_paths(paths, n), std::max<uint32_t>(min_len, 1), one_dataset_per_file != 0);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

// Packed reads -> device batch; input_frag (optional) marks where each
// packed input starts (the plain launch mode scans one input per launch).
static int batch_from_packed(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                             const uint32_t *dataset, const uint64_t *input_frag, size_t n_inputs,
                             mczi_batch **out) {
    return guarded("failed to load batch", [&] {
        require(out != nullptr && frag_off != nullptr, MCZI_E_INVALID, "null argument");
        require(words != nullptr || frag_off[n_frag] == 0, MCZI_E_INVALID, "null argument");
        for (size_t i = 0; i < n_frag; i++)
            require(frag_off[i + 1] >= frag_off[i], MCZI_E_INVALID, "fragment offsets must not decrease");
        if (input_frag) {
            require(n_inputs >= 1 && input_frag[0] == 0 && input_frag[n_inputs] == n_frag, MCZI_E_INVALID,
                    "input marks must start at fragment 0 and end at n_frag");
            for (size_t c = 0; c < n_inputs; c++)
                require(input_frag[c + 1] >= input_frag[c], MCZI_E_INVALID, "input marks must not decrease");
        }
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        uint32_t nd = 1;
        if (dataset)
            for (size_t i = 0; i < n_frag; i++) nd = std::max(nd, dataset[i] + 1);
        // empty fragments are dropped (dataset ids and input marks follow)
        std::vector<uint64_t> c_off, c_inp;
        std::vector<uint32_t> c_ds;
        const std::vector<uint64_t> keep = nonempty_fragments(frag_off, n_frag);
        if (!keep.empty() || (n_frag && frag_off[n_frag] == frag_off[0])) {
            c_off.reserve(keep.size() + 1);
            c_off.push_back(frag_off[0]);
            for (uint64_t i : keep) c_off.push_back(frag_off[i + 1]);
            if (dataset) {
                c_ds.reserve(keep.size());
                for (uint64_t i : keep) c_ds.push_back(dataset[i]);
            }
            if (input_frag) {
                c_inp.resize(n_inputs + 1);
                size_t j = 0;
                for (size_t c = 0; c <= n_inputs; c++) {
                    while (j < keep.size() && keep[j] < input_frag[c]) j++;
                    c_inp[c] = j;
                }
                input_frag = c_inp.data();
            }
            frag_off = c_off.data();
            n_frag = keep.size();
            if (dataset) dataset = c_ds.data();
        }
        // input chunks of >= 128 Mi bases (BATCH_BASES, src/lib.rs:28) for the
        // plain launch mode, each input starting a chunk of its own
        std::vector<uint64_t> chunk_frag{0};
        uint64_t acc = 0;
        size_t c = 1;
        for (size_t i = 0; i < n_frag; i++) {
            acc += frag_off[i + 1] - frag_off[i];
            bool cut = acc >= (128ull << 20);
            if (input_frag) {
                while (c < n_inputs && input_frag[c] <= i + 1) {
                    cut = cut || input_frag[c] == i + 1;
                    c++;
                }
            }
            if (cut && i + 1 < n_frag && chunk_frag.back() != i + 1) {
                chunk_frag.push_back(i + 1);
                acc = 0;
            }
        }
        // input boundaries, empty inputs dropped
        std::vector<uint64_t> inputs{0};
        if (input_frag)
            for (size_t c = 1; c <= n_inputs; c++)
                if (input_frag[c] != inputs.back()) inputs.push_back(input_frag[c]);
        auto h = std::make_unique<mczi_batch>();
        h->b = make_batch_packed(words, frag_off, n_frag, dataset, nd, std::move(chunk_frag), std::move(inputs));
        *out = h.release();
    });
}

int mczi_batch_from_packed(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                           const uint32_t *dataset, mczi_batch **out) {
    return batch_from_packed(words, frag_off, n_frag, dataset, nullptr, 0, out);
}

int mczi_batch_from_packed_inputs(const uint32_t *words, const uint64_t *frag_off, size_t n_frag,
                                  const uint32_t *dataset, const uint64_t *input_frag, size_t n_inputs,
                                  mczi_batch **out) {
    if (input_frag == nullptr)
        return guarded("failed to load batch", [&] { require(false, MCZI_E_INVALID, "null argument"); });
    return batch_from_packed(words, frag_off, n_frag, dataset, input_frag, n_inputs, out);
}

int mczi_batch_info(const mczi_batch *b, uint64_t *n_bases, uint64_t *n_frag, uint32_t *n_datasets) {
    return guarded("failed to query batch", [&] {
        require(b != nullptr, MCZI_E_INVALID, "null batch");
        if (n_bases) *n_bases = b->b->n_bases;
        if (n_frag) *n_frag = b->b->n_frag;
        if (n_datasets) *n_datasets = b->b->n_datasets;
    });
}

void mczi_batch_free(mczi_batch *b) {
    if (!b) return;
    try {
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        delete b;
    } catch (...) {
    }
}

int mczi_batch_count(mczi_batch *b, const mczi_counter_config *cfg, int mode, int launch, int order,
                     mczi_run_stats *stats) {
    return guarded("failed to count batch", [&] {
        require(b != nullptr && cfg != nullptr, MCZI_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        if (is_datasets(mode)) validate_datasets(*cfg);
        else validate_normal(*cfg);
        if (is_datasets(mode) && cfg->threshold >= b->b->n_datasets) {
            b->b->n_out = 0;
            if (stats) *stats = mczi_run_stats{};
            return;
        }
        batch_count(*b->b, *cfg, mode, launch, order, stats);
    });
}

int mczi_batch_autotune(mczi_batch *b, const mczi_counter_config *cfg, int mode, int launch, int order,
                        double *instances_per_partition) {
    return guarded("failed to autotune batch", [&] {
        require(b != nullptr && cfg != nullptr, MCZI_E_INVALID, "null argument");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        if (is_datasets(mode)) validate_datasets(*cfg);
        else validate_normal(*cfg);
        const double t = batch_autotune(*b->b, *cfg, mode, launch, order);
        if (instances_per_partition) *instances_per_partition = t;
    });
}

int mczi_batch_download(mczi_batch *b, mczi_counted_kmer **out, size_t *n_out) {
    return guarded("failed to download results", [&] {
        require(b != nullptr, MCZI_E_INVALID, "null batch");
        require(out && n_out, MCZI_E_INVALID, "null output pointer");
        std::lock_guard<std::mutex> lk(Device::get().call_mutex());
        const size_t n = (size_t)b->b->n_out;
        mczi_counted_kmer *p = (mczi_counted_kmer *)malloc(std::max<size_t>(n, 1) * sizeof(mczi_counted_kmer));
        require(p != nullptr, MCZI_E_NOMEM, "out of host memory");
        try {
            batch_download(*b->b, p);
        } catch (...) {
            free(p);
            throw;
        }
        *out = p;
        *n_out = n;
    });
}

// ---- synthetic inputs -------------------------------------------------------
int mczi_synth_generate(const mczi_synth_spec *spec, mczi_synth_reads *out) {
    return guarded("failed to generate synthetic reads", [&] {
        require(spec && out, MCZI_E_INVALID, "null argument");
        *out = mczi_synth_reads{};
        synth_generate(*spec, *out);
    });
}

int mczi_synth_genome(uint64_t genome_len, uint64_t seed, uint32_t **words) {
    return guarded("failed to generate genome", [&] { synth_genome_words(genome_len, seed, words); });
}

int mczi_synth_reads_from(const uint32_t *const *genome_words, const uint64_t *genome_len, uint32_t n_genomes,
                          uint64_t n_reads, uint32_t read_len, double sub_rate, uint64_t seed,
                          mczi_synth_reads *out) {
    return guarded("failed to generate synthetic reads", [&] {
        require(genome_words && genome_len && out, MCZI_E_INVALID, "null argument");
        *out = mczi_synth_reads{};
        synth_reads_from(genome_words, genome_len, n_genomes, n_reads, read_len, sub_rate, seed, *out);
    });
}

int mczi_synth_write_fasta(const mczi_synth_reads *r, const char *path, int64_t dataset_filter) {
    return guarded("failed to write synthetic FASTA", [&] {
        require(r && path, MCZI_E_INVALID, "null argument");
        synth_write_fasta(*r, path, dataset_filter);
    });
}

void mczi_synth_free(mczi_synth_reads *r) {
    if (!r) return;
    free(r->words);
    free(r->frag_off);
    free(r->dataset);
    *r = mczi_synth_reads{};
}

}  // extern "C"

// Errors raised inside the ZI translation unit are reported through the same
// thread-local text as every other entry point.
extern "C" __attribute__((visibility("hidden"))) void mczi_set_error_from_zi(const char *msg) {
    g_err = msg ? msg : "";
}
