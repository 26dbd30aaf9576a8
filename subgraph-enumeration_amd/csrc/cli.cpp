// graph_mining_system: the reference's CLI (main.cpp, utility/command.cpp:8-55)
// on top of libsge.  Flags: -q pattern, -d data graph, -r result file (stats text
// instead of stdout, main.cpp:29-31), -i exploration-style intersection, -u IEP,
// -t threads (accepted for compatibility; the GPU executor ignores it).
// Extensions: --embeddings <path> writes one canonical embedding per counted
// subgraph (lexicographically smallest image tuple over Aut(P), sorted), which the
// reference never writes; --order a,b,c,... fixes the matching order;
// --bag a,b,c,... materialises those pattern vertices as one bag (its relation
// built into a trie on the GPU before the join; Execution Time includes the
// build, as the reference's includes its bag joins and tries), a second --bag
// adds a second bag; --bag auto takes the first bag of the pattern's
// decomposition that is valid for the order and a second one when a valid pair
// exists;
// --device-build builds the data graph's CSR on the GPU.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "sge.h"

extern "C" int sge_canonicalize_embeddings(const sge_pattern *p, uint32_t *rows, uint64_t n);

static void usage() {
    std::cout << "Usage: graph_mining_system [options]\n"
              << "Options:\n"
              << "  -q <path>        Query graph file path\n"
              << "  -d <path>        Data graph file path\n"
              << "  -r <path>        Result file path (optional, print to stdout if not specified)\n"
              << "  -i               Exploration-style intersection (default: WCOJ)\n"
              << "  -u               Use IEP (default: false)\n"
              << "  -t <number>      Number of threads (accepted; the GPU executor ignores it)\n"
              << "  --embeddings <path>  Write canonical embeddings (extension)\n"
              << "  --order a,b,...  Fixed matching order (extension)\n"
              << "  --bag a,b,...|auto  Materialise these pattern vertices (or the first valid\n"
              << "                   decomposition bags) as a bag; twice: two bags (extension)\n"
              << "  --device-build   Build the data graph's CSR on the GPU (extension)\n"
              << "  --help, -h       Show this help message\n"
              << std::endl;
}

static int die(const char *what) {
    std::fprintf(stderr, "error: %s: %s\n", what, sge_last_error());
    return 1;
}

int main(int argc, char **argv) {
    std::string q, d, r, emb, order_s;
    std::vector<std::string> bag_s;
    bool explore = false, iep = false;
    uint32_t gflags = SGE_GRAPH_DEGREE_ORDER;
    int threads = 1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string { return (i + 1 < argc) ? std::string(argv[++i]) : std::string(); };
        if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (a == "-q") q = next();
        else if (a == "-d") d = next();
        else if (a == "-r") r = next();
        else if (a == "-i") explore = true;
        else if (a == "-u") iep = true;
        else if (a == "-t") threads = std::atoi(next().c_str());
        else if (a == "--embeddings") emb = next();
        else if (a == "--order") order_s = next();
        else if (a == "--bag") bag_s.push_back(next());
        else if (a == "--device-build") gflags |= SGE_GRAPH_BUILD_DEVICE;
    }
    if (threads <= 0) threads = 1;  // atoi of a missing value -> 0 -> 1 (main.cpp:121)
    std::ofstream fout;
    std::ostream *out = &std::cout;
    if (!r.empty()) {
        fout.open(r);
        if (!fout) {
            std::fprintf(stderr, "error: can not create result file %s\n", r.c_str());
            return 1;
        }
        out = &fout;
    }
    sge_pattern p;
    if (sge_pattern_load(q.c_str(), &p)) return die("pattern");
    sge_graph *g = nullptr;
    if (sge_graph_load(d.c_str(), gflags, &g)) return die("data graph");
    auto t0 = std::chrono::steady_clock::now();
    uint32_t mode = iep ? SGE_MODE_IEP : (explore ? SGE_MODE_EXPLORE : SGE_MODE_WCOJ);
    std::vector<uint32_t> order;
    if (!order_s.empty()) {
        std::stringstream ss(order_s);
        std::string tok;
        while (std::getline(ss, tok, ',')) order.push_back((uint32_t)std::atoi(tok.c_str()));
        if (order.size() != p.k) {
            std::fprintf(stderr, "error: --order needs %u vertices\n", p.k);
            return 1;
        }
    }
    sge_plan plan;
    if (sge_plan_build(&p, g, mode, order.empty() ? nullptr : order.data(), &plan)) return die("plan");
    if (bag_s.size() > 2 || (bag_s.size() == 2 && (bag_s[0] == "auto" || bag_s[1] == "auto"))) {
        std::fprintf(stderr, "error: --bag: at most two bags, or one --bag auto\n");
        return 1;
    }
    if (bag_s.size() == 1 && bag_s[0] == "auto") {
        uint32_t bags[16], nb = 0;
        if (sge_plan_bags(&p, &plan, bags, 16, &nb)) return die("bag");
        if (nb && sge_plan_set_bag(&plan, bags[0])) return die("bag");
        for (uint32_t i = 1; i < nb && i < 16; ++i) {
            sge_plan two = plan;
            if (sge_plan_add_bag(&two, bags[i]) == SGE_OK) {
                plan = two;
                break;
            }
        }
    } else {
        for (size_t b = 0; b < bag_s.size(); ++b) {
            std::stringstream ss(bag_s[b]);
            std::string tok;
            uint32_t mask = 0;
            while (std::getline(ss, tok, ',')) {
                char *end = nullptr;
                const unsigned long v = std::strtoul(tok.c_str(), &end, 10);
                if (tok.empty() || *end || v >= p.k) {
                    std::fprintf(stderr, "error: --bag: '%s' is not a pattern vertex (0..%u)\n", tok.c_str(),
                                 p.k - 1);
                    return 1;
                }
                mask |= 1u << v;
            }
            if ((b == 0 ? sge_plan_set_bag(&plan, mask) : sge_plan_add_bag(&plan, mask))) return die("bag");
        }
    }
    double plan_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    char text[4096];
    if (sge_plan_text(&plan, text, sizeof(text)) == 0) *out << text;
    *out << "Planning time (s): " << plan_s << std::endl;
    *out << "traverse all level: 0" << std::endl;
    *out << "use iep: " << (iep ? 1 : 0) << std::endl;
    if (threads > 1) *out << "Number of threads: " << threads << std::endl;
    if (iep && plan.tail >= 2) *out << "optimized traversal" << std::endl;
    uint64_t count = 0;
    sge_stats st;
    std::memset(&st, 0, sizeof(st));
    if (sge_count(g, &plan, 1, 0, nullptr, &count, &st, nullptr)) return die("count");
    double exec_s = (st.kernel_ms + st.trie_ms) / 1000.0;
    *out << "Number of matches: " << count << std::endl;
    if (threads > 1) {
        *out << "Execution Time: " << exec_s << std::endl;
        *out << "Number of intersections: " << st.intersections << std::endl;
    } else {
        *out << "Number of intersections: " << st.intersections << std::endl;
        *out << "Execution Time: " << exec_s << std::endl;
    }
    if (!emb.empty()) {
        uint64_t n = 0;
        uint64_t cap = count / (plan.multiplier ? plan.multiplier : 1) + 16;
        if (cap > (1ull << 32)) {  // the host copy alone would exceed 128 GiB
            std::fprintf(stderr, "error: --embeddings: %llu embeddings are too many to write\n",
                         (unsigned long long)cap);
            return 1;
        }
        std::vector<uint32_t> rows((size_t)cap * p.k);
        int rc = sge_enumerate(g, &plan, rows.data(), cap, &n);
        if (rc) return die("enumerate");
        if (sge_canonicalize_embeddings(&p, rows.data(), n)) return die("embeddings");
        std::ofstream ef(emb);
        if (!ef) {
            std::fprintf(stderr, "error: can not create embeddings file %s\n", emb.c_str());
            return 1;
        }
        ef << n << "\n";
        for (uint64_t i = 0; i < n; ++i) {
            for (uint32_t j = 0; j < p.k; ++j) ef << rows[i * p.k + j] << (j + 1 < p.k ? " " : "");
            ef << "\n";
        }
    }
    sge_graph_destroy(g);
    return 0;
}
