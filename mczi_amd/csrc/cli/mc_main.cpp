// MC command line (src/main.rs:10-103): same flags, same dispatch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/mczi.h"

static void usage(FILE *f) {
    fprintf(f,
            "Parallel two-pass pseudo k-mer counter using minimizer pre-counting\n\n"
            "Usage: MC [OPTIONS] --input <INPUT>... --kmer-size <KMER_SIZE> --minimizer-size "
            "<MINIMIZER_SIZE> --output <OUTPUT>\n\n"
            "Options:\n"
            "  -i, --input <INPUT>...                   Input FASTA/FASTQ file(s)\n"
            "      --fofn                               Treat --input path(s) as file-of-filenames; -x is a "
            "dataset-presence threshold\n"
            "  -k, --kmer-size <KMER_SIZE>              K-mer size; must be <= 64\n"
            "  -m, --minimizer-size <MINIMIZER_SIZE>    Minimizer size; must be <= k\n"
            "  -x, --threshold <THRESHOLD>              Emit k-mers seen more than this many times [default: 1]\n"
            "  -o, --output <OUTPUT>                    Output FASTA or KFF file\n"
            "      --format <FORMAT>                    Output format [default: fasta] [possible values: fasta, kff]\n"
            "  -t, --threads <THREADS>                  Number of host worker threads\n"
            "      --partition-count <PARTITION_COUNT>  Number of partitions used by counting phases [default: 1024]\n"
            "  -h, --help                               Print help\n");
}

static int fail(const char *msg) {
    fprintf(stderr, "Error: %s\n", msg);
    return 1;
}

int main(int argc, char **argv) {
    std::vector<std::string> inputs;
    bool fofn = false, have_k = false, have_m = false, have_o = false;
    unsigned long long k = 0, m = 0, x = 1, parts = MCZI_DEFAULT_PARTITION_COUNT, threads = 0;
    std::string out, format = "fasta";
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto val = [&](const char *name) -> std::string {
            size_t eq = a.find('=');
            if (a.rfind("--", 0) == 0 && eq != std::string::npos) return a.substr(eq + 1);
            if (i + 1 >= argc) {
                fprintf(stderr, "error: a value is required for '%s'\n", name);
                exit(2);
            }
            return argv[++i];
        };
        auto num = [&](const char *name) -> unsigned long long {
            std::string v = val(name);
            char *end = nullptr;
            unsigned long long r = strtoull(v.c_str(), &end, 10);
            if (v.empty() || *end) {
                fprintf(stderr, "error: invalid value '%s' for '%s'\n", v.c_str(), name);
                exit(2);
            }
            return r;
        };
        auto is = [&](const char *s, const char *l) {
            return a == s || a == l || (a.rfind(std::string(l) + "=", 0) == 0);
        };
        if (a == "-h" || a == "--help") {
            usage(stdout);
            return 0;
        } else if (is("-i", "--input")) {
            inputs.push_back(val("--input"));
            while (i + 1 < argc && argv[i + 1][0] != '-') inputs.push_back(argv[++i]);
        } else if (a == "--fofn") {
            fofn = true;
        } else if (is("-k", "--kmer-size")) {
            k = num("--kmer-size");
            have_k = true;
        } else if (is("-m", "--minimizer-size")) {
            m = num("--minimizer-size");
            have_m = true;
        } else if (is("-x", "--threshold")) {
            x = num("--threshold");
        } else if (is("-o", "--output")) {
            out = val("--output");
            have_o = true;
        } else if (is("--format", "--format")) {
            format = val("--format");
            if (format != "fasta" && format != "kff") {
                fprintf(stderr, "error: invalid value '%s' for '--format'\n", format.c_str());
                return 2;
            }
        } else if (is("-t", "--threads")) {
            threads = num("--threads");
            if (threads == 0) return fail("--threads must be greater than 0");
        } else if (is("--partition-count", "--partition-count")) {
            parts = num("--partition-count");
        } else {
            fprintf(stderr, "error: unexpected argument '%s'\n", a.c_str());
            usage(stderr);
            return 2;
        }
    }
    if (inputs.empty() || !have_k || !have_m || !have_o) {
        fprintf(stderr, "error: the following required arguments were not provided: --input --kmer-size "
                        "--minimizer-size --output\n");
        return 2;
    }
    mczi_counter_config cfg{(uint32_t)k, (uint32_t)m, x, (uint32_t)parts};
    if (k > 0xffffffffull || m > 0xffffffffull || parts > 0xffffffffull) return fail("value out of range");
    // the global worker pool (src/main.rs:73-79)
    if (threads) mczi_set_threads((uint32_t)threads);
    std::vector<const char *> p;
    for (auto &s : inputs) p.push_back(s.c_str());
    int rc;
    if (fofn) {
        char **ds = nullptr;
        size_t nd = 0;
        rc = mczi_expand_fofns(p.data(), p.size(), &ds, &nd);
        if (rc) return fail(mczi_last_error());
        rc = format == "fasta" ? mczi_count_datasets_to_fasta_path(ds, nd, &cfg, out.c_str())
                               : mczi_count_datasets_to_kff_path(ds, nd, &cfg, out.c_str());
        mczi_free_strings(ds, nd);
    } else {
        rc = format == "fasta" ? mczi_count_inputs_to_fasta_path(p.data(), p.size(), &cfg, out.c_str())
                               : mczi_count_inputs_to_kff_path(p.data(), p.size(), &cfg, out.c_str());
    }
    if (rc) return fail(mczi_last_error());
    return 0;
}
