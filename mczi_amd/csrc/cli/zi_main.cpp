// ZI command line (src/zi.rs:13-100): same flags and checks.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/mczi.h"

static int fail(const char *msg) {
    fprintf(stderr, "Error: %s\n", msg);
    return 1;
}

int main(int argc, char **argv) {
    std::vector<std::string> inputs;
    std::string index, out;
    bool fofn = false, have_k = false, have_m = false;
    unsigned long long k = 0, m = 0, threads = 0, ram = 8;
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
            printf("Subtract an indexed k-mer set from FASTA/FOFN FASTA input\n\n"
                   "Usage: ZI [OPTIONS] --index <INDEX> --input <INPUT>... --kmer-size <KMER_SIZE> --output <OUTPUT>\n"
                   "  --index <INDEX>  -i/--input <INPUT>...  --fofn  -k/--kmer-size  -m/--minimizer-size\n"
                   "  -o/--output  -t/--threads [default: 0]  --ram-limit-gib [default: 8]\n");
            return 0;
        } else if (is("--index", "--index")) {
            index = val("--index");
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
        } else if (is("-o", "--output")) {
            out = val("--output");
        } else if (is("-t", "--threads")) {
            threads = num("--threads");
        } else if (is("--ram-limit-gib", "--ram-limit-gib")) {
            ram = num("--ram-limit-gib");
        } else {
            fprintf(stderr, "error: unexpected argument '%s'\n", a.c_str());
            return 2;
        }
    }
    if (index.empty() || inputs.empty() || !have_k || out.empty()) {
        fprintf(stderr, "error: the following required arguments were not provided: --index --input "
                        "--kmer-size --output\n");
        return 2;
    }
    // validate_k (src/zi.rs:103-107) and the minimizer check (src/zi.rs:70-74)
    if (!(k >= 3 && k <= 63)) return fail("k must be in [3, 63] for sshash-rs");
    if (k % 2 == 0) return fail("k must be odd for sshash-rs");
    if (!have_m) {
        m = k - 2 < 19 ? k - 2 : 19;
        if (m % 2 == 0) m -= 1;
    }
    if (!(m > 0 && m < k)) return fail("minimizer size must be > 0 and < k");
    std::vector<const char *> p;
    for (auto &s : inputs) p.push_back(s.c_str());
    char **qs = nullptr;
    size_t nq = 0;
    const char *const *qp = p.data();
    if (fofn) {
        if (mczi_expand_fofns(p.data(), p.size(), &qs, &nq)) return fail(mczi_last_error());
        qp = qs;
    } else {
        nq = p.size();
    }
    if (nq == 0) return fail("at least one query FASTA is required");
    int rc = mczi_zi_subtract(index.c_str(), qp, nq, (uint32_t)k, (uint32_t)m, out.c_str(), (uint32_t)threads,
                              (uint32_t)ram);
    if (qs) mczi_free_strings(qs, nq);
    if (rc) return fail(mczi_last_error());
    return 0;
}
