// pgd_main.cpp -- the `pgd` driver: same flags and output sequence as the
// reference driver (graphlet_driver.cpp:53-92, graphlet_params.h:216-287),
// with the decomposition on the GPU.
#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "graphlet_core.hpp"

namespace pgd_amd {

void params::print_help(const char* argv0) {
    std::fprintf(stderr,
                 "Usage: %s -f path [options]\n"
                 "  Exact 2-, 3- and 4-node graphlet counts (global and per edge) on an AMD MI355X.\n"
                 "  -f, --file, --graph, -g FILE   input graph (.mtx, or an edge list: .txt .edges .csv ...)\n"
                 "  -c, --counts, --macro FILE     write the 17 global counts\n"
                 "  -m, --micro FILE               write the per-edge counts (one row per edge)\n"
                 "  -o, --ordering, --order NAME   edge order (deg, kcore, natural, kcore_deg, deg_vol, ...)\n"
                 "      --s2l, --small_to_large    order edges from small to large\n"
                 "  -n, --neigh_ordering NAME      neighbour order (deg, kcore, ...)\n"
                 "      --s2l_neigh                order neighbours from small to large\n"
                 "  -w, --workers N  -b, --block_size N  -r, --rep NAME  -l, --adj_limit N\n"
                 "                                 accepted for compatibility; the device schedules the work\n"
                 "  -a, --algorithm NAME           exact (the only algorithm)\n"
                 "  -v, --verbose  -h, --help\n",
                 argv0);
}

// flag tables and parsing rules of graphlet_params.h:216-287: boolean flags
// take no value; other flags consume the next argument when there is one;
// unknown flags are reported and ignored
params::params(int argc, char** argv) {
    if (argc <= 1) { print_help(argv[0]); std::exit(-1); }
    std::map<std::string, int*> ints = {{"-w", &workers}, {"--workers", &workers}, {"--threads", &workers},
                                        {"--cores", &workers}, {"-b", &block_size}, {"--block_size", &block_size},
                                        {"-l", &adj_limit}, {"--adj_limit", &adj_limit}, {"--adj_size", &adj_limit}};
    std::map<std::string, bool*> bools = {{"-?", &help}, {"-h", &help}, {"--help", &help}, {"-v", &verbose},
                                          {"--verbose", &verbose}, {"--s2l", &is_small_to_large},
                                          {"--small_to_large", &is_small_to_large},
                                          {"--s2l_neigh", &is_small_to_large_csc_neighbor},
                                          {"--small_to_large_neigh", &is_small_to_large_csc_neighbor}};
    std::map<std::string, std::string*> strs = {
        {"-f", &graph}, {"--graph", &graph}, {"--file", &graph}, {"-g", &graph}, {"-a", &algorithm},
        {"--algorithm", &algorithm}, {"--alg", &algorithm}, {"-c", &macro_stats_filename},
        {"--counts", &macro_stats_filename}, {"--macro", &macro_stats_filename}, {"-m", &micro_stats_filename},
        {"--micro", &micro_stats_filename}, {"-r", &graph_representation}, {"--rep", &graph_representation},
        {"-o", &ordering}, {"--ordering", &ordering}, {"--order", &ordering}, {"-n", &ordering_csc_neighbor},
        {"--neigh_ordering", &ordering_csc_neighbor}, {"--neigh_order", &ordering_csc_neighbor},
        {"--csc_ordering", &ordering_csc_neighbor}, {"--neighbor_ordering", &ordering_csc_neighbor}};
    for (int i = 1; i < argc; ++i) {
        std::string flag(argv[i]);
        auto b = bools.find(flag);
        if (b != bools.end()) { *b->second = true; continue; }
        if (i + 1 < argc) {
            auto it = ints.find(flag);
            auto st = strs.find(flag);
            if (it != ints.end()) *it->second = std::atoi(argv[++i]);
            else if (st != strs.end()) *st->second = std::string(argv[++i]);
            else std::printf("Invalid argument %s, ignoring.\n", flag.c_str());
        }
    }
    if (help) { print_help(argv[0]); std::exit(-1); }
    if (workers <= 0) workers = 1;
    std::ifstream probe(graph.c_str());
    if (!probe) { print_help(argv[0]); std::exit(-1); }
}

}  // namespace pgd_amd

namespace {

// univariate statistics of a per-edge column, with the reference's
// arithmetic (graphlet_univar_stats.h:70-110: integer mean and median)
std::string univar(std::vector<unsigned long long> data) {
    double sum_sq = 0;
    unsigned long long sum = 0, mx = 0, mn = 0, cnt = 0;
    for (unsigned long long x : data) {
        sum_sq += (double)(x * x);
        sum += x;
        if (x > mx) mx = x;
        if (x < mn) mn = x;
        cnt++;
    }
    std::ostringstream os;
    if (cnt == 0) return os.str();
    double mean = (double)(sum / cnt);
    double var = (sum_sq / cnt) - mean * mean;
    double sd = std::sqrt(var);
    unsigned long long range = mx - mn;
    std::sort(data.begin(), data.end());
    unsigned long long med_idx = cnt / 2, qi = cnt / 4;
    double median = (double)data[med_idx];
    if (data.size() % 2 == 0) median = (double)((data[med_idx - 1] + data[med_idx]) / 2);
    double q1 = (double)data[med_idx - qi], q3 = (double)data[med_idx + qi];
    double iqr = q3 - q1;
    os << "mean\t= " << mean << "\n" << "median\t= " << median << "\n" << "max\t= " << mx << "\n"
       << "min\t= " << mn << "\n" << "range\t= " << range << "\n" << "std\t= " << sd << "\n"
       << "var\t= " << var << "\n" << "iqr\t= " << iqr << "\n" << "q1\t= " << q1 << "\n" << "q3\t= " << q3
       << "\n";
    return os.str();
}

double now() {
    struct timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec + t.tv_usec / 1e6;
}

}  // namespace

int main(int argc, char* argv[]) {
    using namespace pgd_amd;
    try {
        params p(argc, argv);
        if (p.verbose) std::cout << "graph file: " << p.graph << std::endl;
        graphlet_core G(p);
        G.compute_assortativity();
        std::cout << "r = " << G.get_assortativity() << std::endl;
        G.compute_cores();
        std::cout << "K = " << G.get_max_core() << std::endl;
        double s = now();
        G.sort_edges(p.ordering, p.is_small_to_large);
        if (p.is_micro_stats()) G.graphlet_decomposition_micro(p.workers);
        else G.graphlet_decomposition(p.workers);
        s = now() - s;
        G.print_graphlet_counts();
        if (p.is_micro_stats()) G.print_micro_stats();
        std::cout << "graphlet decomposition time: " << s << " sec" << std::endl;
        if (p.is_macro_stats()) G.write_macro_stats(p.macro_stats_filename);
        if (p.is_micro_stats()) G.write_micro_stats(p.micro_stats_filename);
        G.print_GFD();
        G.print_connected_GFD();
        G.print_disconnected_GFD();
        if (p.is_micro_stats()) std::cout << univar(G.local_4_clique) << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "pgd: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
