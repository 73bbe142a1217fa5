// graphlet_core.hpp -- host C++ mirror of graphlet::graphlet_core over the
// C ABI (include/pgd_hip.h): same member names, argument meaning and public
// result fields as the reference class (graphlet_core.h:54-703) for the
// decomposition path, with the counting done on the GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "pgd_hip.h"

namespace pgd_amd {

// Command-line parameters (graphlet_params.h:41-317): same flags and defaults.
struct params {
    std::string algorithm = "exact";
    int adj_limit = 10000;
    bool graph_stats = false, verbose = false, help = false;
    std::string graph = "sample_graph.csv";
    std::string macro_stats_filename, micro_stats_filename;
    std::string ordering = "deg";
    bool is_small_to_large = false;
    std::string ordering_csc_neighbor = "deg";
    bool is_small_to_large_csc_neighbor = false;
    std::string graph_representation;
    int workers = 1;
    int block_size = 64;
    params() = default;
    params(int argc, char** argv);
    bool is_micro_stats() const { return !micro_stats_filename.empty(); }
    bool is_macro_stats() const { return !macro_stats_filename.empty(); }
    static void print_help(const char* argv0);
};

class graphlet_core {
public:
    explicit graphlet_core(params& p);
    graphlet_core(const std::string& filename, std::string ordering_technique = "degree",
                  bool is_small_to_large = false, int adj_limit = 10000, std::string rep = "auto",
                  int block_sz = 64);
    ~graphlet_core();
    graphlet_core(const graphlet_core&) = delete;
    graphlet_core& operator=(const graphlet_core&) = delete;

    void preprocess(std::string ordering_csc_neighbor = "degree", bool is_small_to_large = false,
                    int adj_limit = 10000, std::string rep = "auto");
    void compute_cores();
    void compute_assortativity();
    // sort_edges appends to E_ordered, as the reference does (graphlet_core.cpp:854-867)
    void sort_edges(std::string edge_ordering = "degree", bool is_small_to_large = false);
    // the overload over a sampled edge set E_s (graphlet_core.cpp:836-851)
    void sort_edges(std::string ordering_technique, bool is_small_to_large, std::vector<unsigned long>& E_s);
    void graphlet_decomposition(int max_num_workers = 1);
    void graphlet_decomposition_micro(int max_num_workers = 1);

    int num_vertices() const { return (int)n_; }
    int num_edges() const { return (int)m_; }
    double density() const { return (double)num_edges() / (num_vertices() * (num_vertices() - 1.0) / 2.0); }
    int get_max_degree() const { return max_degree; }
    double get_avg_degree() const { return avg_degree; }
    int get_max_core() const { return max_core; }
    double get_assortativity() const { return r; }

    void basic_stats(std::string prefix = "", std::string suffix = "\n");
    void print_graphlet_counts();
    std::string get_graphlet_counts();
    void write_macro_stats(std::string filename);
    std::string get_micro_stats(bool output_id = true, std::string delim = ",");
    void print_micro_stats(bool output_id = true, std::string delim = ",");
    void write_micro_stats(std::string& filename, bool output_id = true);
    std::string compute_GFD();
    std::string compute_connected_GFD();
    std::string compute_disconnected_GFD();
    void print_GFD();
    void print_connected_GFD();
    void print_disconnected_GFD();

    // results (graphlet_core.h:310-343); E_ordered is public as in the
    // reference (graphlet_core.h:207): the decomposition runs over whatever
    // edge ids it holds
    unsigned long long total_2_1edge = 0, total_2_indep = 0, total_3_tris = 0, total_2_star = 0, total_3_1edge = 0,
                       total_3_indep = 0, total_4_clique = 0, total_4_chordcycle = 0, total_4_tailed_tris = 0,
                       total_4_cycle = 0, total_3_star = 0, total_4_path = 0, total_4_2edge = 0, total_4_1edge = 0,
                       total_4_2star = 0, total_4_tri = 0, total_4_indep = 0;
    std::vector<unsigned long long> tri, local_4_clique, local_4_chordal_cycle, local_4_tailed_tris, local_4_cycle,
        local_3_star, local_4_path, two_star;
    std::vector<int32_t> e_v, e_u, E_ordered;
    std::vector<int32_t> degree;
    int max_degree = 0, avg_degree = 0, max_core = 0;
    double r = 0;
    double decomposition_ms = 0;  // device time of the last decomposition's counting
    std::string fn;

private:
    void read_graph(const std::string& filename);
    void decompose(bool micro);
    void fetch_edges();
    void set_totals(const uint64_t* t);
    pgd_graph* g_ = nullptr;
    void append_order();
    int64_t n_ = 0, m_ = 0;
    uint32_t engine_ = 0;  // PGD_ENGINE_* from the representation (-r), 0: the library's choice
};

// pgd_ctx shared by every graphlet_core of the process (device 0)
pgd_ctx* default_context();
void check(pgd_status st, const char* what);

}  // namespace pgd_amd
