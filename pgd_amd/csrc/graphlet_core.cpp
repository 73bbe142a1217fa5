// graphlet_core.cpp -- host mirror of graphlet::graphlet_core over the C ABI.
//
// Reading, preprocessing, k-core, edge order and counting are C-ABI calls
// (the device does the work); this file keeps the reference's object model
// and its text outputs: the count block (graphlet_core.cpp:2061-2084), the
// .macro file (2138-2162), the micro dump and .micro file (1668-1828),
// basic_stats (562-568), assortativity (1029-1057) and the GFD
// (1955-2056).
#include "graphlet_core.hpp"

#include <algorithm>
#include <fcntl.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace pgd_amd {

namespace {

const char* kTotalNames[PGD_NTOTALS] = {
    "total_2_1edge", "total_2_indep", "total_3_tris", "total_2_star", "total_3_1edge", "total_3_indep",
    "total_4_clique", "total_4_chordcycle", "total_4_tailed_tris", "total_4_cycle", "total_3_star",
    "total_4_path", "total_4_1edge", "total_4_2edge", "total_4_2star", "total_4_tri", "total_4_indep"};

void print_line(int n = 80, const std::string& sym = "-") {
    for (int i = 0; i < n; ++i) std::cout << sym;
    std::cout << std::endl;
}

std::string extension(const std::string& f) {
    size_t d = f.rfind('.');
    return d == std::string::npos ? std::string() : f.substr(d + 1);
}

// decimal formatting without iostream overhead for the large micro tables
inline char* put_u64(char* p, unsigned long long v) {
    char tmp[24];
    int k = 0;
    do { tmp[k++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (k) *p++ = tmp[--k];
    return p;
}

}  // namespace

void check(pgd_status st, const char* what) {
    if (st != PGD_OK) {
        std::string msg = std::string(what) + ": " + pgd_status_str(st);
        pgd_ctx* c = default_context();
        if (c && *pgd_last_error(c)) msg += std::string(" (") + pgd_last_error(c) + ")";
        throw std::runtime_error(msg);
    }
}

pgd_ctx* default_context() {
    static pgd_ctx* ctx = nullptr;
    static bool tried = false;
    if (!tried) {
        tried = true;
        pgd_status st = pgd_open(0, &ctx);
        if (st != PGD_OK) throw std::runtime_error(std::string("pgd_open: ") + pgd_status_str(st));
    }
    return ctx;
}

// ---------------------------------------------------------------------------
graphlet_core::graphlet_core(params& p) {
    fn = p.graph;
    read_graph(p.graph);
    preprocess(p.ordering_csc_neighbor, p.is_small_to_large_csc_neighbor, p.adj_limit, p.graph_representation);
}

graphlet_core::graphlet_core(const std::string& filename, std::string ordering_technique, bool is_small_to_large,
                             int adj_limit, std::string rep, int /*block_sz*/) {
    fn = filename;
    read_graph(filename);
    preprocess(ordering_technique, is_small_to_large, adj_limit, rep);
}

graphlet_core::~graphlet_core() {
    if (g_) pgd_graph_free(g_);
}

// read_graph (graphlet_core.cpp:259-283): parse, build the CSR on the device,
// degrees, basic stats
void graphlet_core::read_graph(const std::string& filename) {
    std::string ext = extension(filename);
    // the reader's informational lines (graphlet_core.cpp:297, 207, 347, 479)
    {
        std::ifstream f(filename.c_str());
        std::string line;
        bool got = false;
        while (std::getline(f, line)) {
            if (!line.empty() && (line[0] == '#' || line[0] == '%')) continue;
            got = true;
            break;
        }
        if (!got) line.clear();
        if (ext != "mtx") {
            std::cout << "[reading generic edge list: read_edge_list func]  filename: " << filename << std::endl;
            bool header_pattern = false;
            {
                std::ifstream g(filename.c_str());
                std::string l;
                while (std::getline(g, l) && !l.empty() && (l[0] == '#' || l[0] == '%'))
                    if (l.find("MatrixMarket matrix coordinate pattern symmetric") != std::string::npos)
                        header_pattern = true;
            }
            if (!header_pattern && line.find(',') == std::string::npos && line.find('\t') == std::string::npos &&
                line.find("   ") != std::string::npos)
                std::printf("found tab-spaces delimiter \n");
            std::istringstream ss(line);
            std::string tok;
            int ntok = 0;
            while (ss >> tok) ntok++;
            if (ntok == 3) std::printf("weighted graph detected \n");
            std::fflush(stdout);
        } else {
            std::istringstream ss(line);
            long long rows = 0, cols = 0;
            ss >> rows >> cols;
            if (rows != cols) std::cout << "[pgd]  error: this is not a square matrix, attempting to proceed." << std::endl;
        }
    }
    // the text parsed by all host threads; renumbering, dedup and CSR on the device
    int64_t loops = 0;
    check(pgd_graph_from_file(default_context(), filename.c_str(), 0, &g_, &loops), "read_graph");
    int32_t md = 0, ad = 0;
    check(pgd_graph_info(g_, &n_, &m_, &md, &ad), "pgd_graph_info");
    max_degree = md;
    avg_degree = ad;
    std::vector<int64_t> row((size_t)n_ + 1);
    std::vector<int32_t> col((size_t)std::max<int64_t>(2 * m_, 1));
    check(pgd_graph_csr(g_, row.data(), col.data()), "pgd_graph_csr");
    degree.resize((size_t)n_);
    for (int64_t v = 0; v < n_; v++) degree[(size_t)v] = (int32_t)(row[(size_t)v + 1] - row[(size_t)v]);
    basic_stats();
}

// The representation (-r) as the reference applies it (optimize_graph_ops and
// set_rep_manually, graphlet_core.cpp:602-611, 663-684), mapped onto the two
// engines, which give the same counts: csc -> the degree-oriented engine;
// adj and hybrid, where the reference builds its adjacency matrix -> the
// dense tier (the oriented engine above PGD_DENSE_MAX_N vertices, which the
// tier does not hold); auto (or none) -> the library's choice, by speed.
void graphlet_core::preprocess(std::string neigh, bool s2l, int /*adj_limit*/, std::string rep) {
    check(pgd_preprocess(g_, neigh.c_str(), s2l ? 1 : 0), "preprocess");
    if (rep == "csc") engine_ = PGD_ENGINE_ORIENT;
    else if (rep == "adj" || rep == "hybrid") engine_ = n_ <= PGD_DENSE_MAX_N ? PGD_ENGINE_DENSE : PGD_ENGINE_ORIENT;
    else engine_ = 0;
    fetch_edges();
}

void graphlet_core::fetch_edges() {
    e_v.resize((size_t)m_);
    e_u.resize((size_t)m_);
    if (m_) check(pgd_edge_arrays(g_, e_v.data(), e_u.data()), "pgd_edge_arrays");
}

void graphlet_core::compute_cores() {
    int32_t K = 0;
    check(pgd_compute_cores(g_, nullptr, &K), "compute_cores");
    max_core = K;
}

// compute_assortativity (graphlet_core.cpp:1029-1057): int products as the
// reference computes them (they wrap for degrees above 46340)
void graphlet_core::compute_assortativity() {
    double mu = 0, jd = 0, psi = 0, tau;
    for (int64_t e = 0; e < m_; e++) {
        int dv = degree[(size_t)e_v[(size_t)e]], du = degree[(size_t)e_u[(size_t)e]];
        mu += double((int)((unsigned)dv * (unsigned)du));
        jd += 0.5 * double(dv + du);
        psi += 0.5 * double((int)((unsigned)dv * (unsigned)dv + (unsigned)du * (unsigned)du));
    }
    double m = (double)m_;
    tau = (jd / m) * (jd / m);
    r = ((mu / m) - tau) / ((psi / m) - tau);
    if (std::isnan(r)) r = 1;
    if (r < -1 || r > 1.0) r = 1;
    if (m == 0) r = 0.0;
}

// the device's new order appended to E_ordered
void graphlet_core::append_order() {
    int64_t k = 0;
    check(pgd_e_ordered_count(g_, &k), "pgd_e_ordered_count");
    const size_t at = E_ordered.size();
    E_ordered.resize(at + (size_t)k);
    if (k) check(pgd_e_ordered(g_, E_ordered.data() + at), "pgd_e_ordered");
}

void graphlet_core::sort_edges(std::string ordering, bool s2l) {
    check(pgd_sort_edges(g_, ordering.c_str(), s2l ? 1 : 0), "sort_edges");
    append_order();
}

void graphlet_core::sort_edges(std::string ordering, bool s2l, std::vector<unsigned long>& E_s) {
    std::vector<int32_t> ids(E_s.size());
    for (size_t i = 0; i < E_s.size(); i++) ids[i] = (int32_t)E_s[i];  // `int edge_id = E_s[e]` (843)
    check(pgd_sort_edges_subset(g_, ordering.c_str(), s2l ? 1 : 0, ids.data(), (int64_t)ids.size()),
          "sort_edges");
    append_order();
}

void graphlet_core::set_totals(const uint64_t* t) {
    unsigned long long* f[PGD_NTOTALS] = {
        &total_2_1edge, &total_2_indep, &total_3_tris, &total_2_star, &total_3_1edge, &total_3_indep,
        &total_4_clique, &total_4_chordcycle, &total_4_tailed_tris, &total_4_cycle, &total_3_star,
        &total_4_path, &total_4_1edge, &total_4_2edge, &total_4_2star, &total_4_tri, &total_4_indep};
    for (int i = 0; i < PGD_NTOTALS; i++) *f[i] = t[i];
}

void graphlet_core::decompose(bool micro) {
    if (E_ordered.empty()) sort_edges();  // graphlet_core.cpp:1376-1378
    check(pgd_set_e_ordered(g_, E_ordered.data(), (int64_t)E_ordered.size()), "E_ordered");
    uint64_t sums[PGD_NSUMS], tot[PGD_NTOTALS];
    check(pgd_decompose(g_, (micro ? PGD_MICRO : 0u) | engine_, sums), "graphlet_decomposition");
    check(pgd_derive(n_, m_, sums, tot), "pgd_derive");
    set_totals(tot);
    double total = 0;
    pgd_last_timing(g_, &total, nullptr, nullptr);
    decomposition_ms = total;
    if (micro) {
        std::vector<unsigned long long>* cols[PGD_NMICRO] = {&tri, &local_4_clique, &local_4_chordal_cycle,
                                                             &local_4_tailed_tris, &local_4_cycle, &local_3_star,
                                                             &local_4_path, &two_star};
        for (int c = 0; c < PGD_NMICRO; c++) {
            cols[c]->resize((size_t)m_);
            if (m_) check(pgd_micro(g_, c, (uint64_t*)cols[c]->data()), "pgd_micro");
        }
    }
}

void graphlet_core::graphlet_decomposition(int) { decompose(false); }
void graphlet_core::graphlet_decomposition_micro(int) { decompose(true); }

// ---------------------------------------------------------------------------
void graphlet_core::basic_stats(std::string prefix, std::string suffix) {
    std::cout << prefix << "|V|: " << num_vertices() << suffix;
    std::cout << prefix << "|E|: " << num_edges() << suffix;
    std::cout << prefix << "p: " << density() << suffix;
    std::cout << prefix << "d_max: " << get_max_degree() << suffix;
    std::cout << prefix << "d_avg: " << get_avg_degree() << suffix;
}

void graphlet_core::print_graphlet_counts() {
    const unsigned long long v[PGD_NTOTALS] = {
        total_2_1edge, total_2_indep, total_3_tris, total_2_star, total_3_1edge, total_3_indep,
        total_4_clique, total_4_chordcycle, total_4_tailed_tris, total_4_cycle, total_3_star,
        total_4_path, total_4_1edge, total_4_2edge, total_4_2star, total_4_tri, total_4_indep};
    print_line(60, "*");
    for (int i = 0; i < PGD_NTOTALS; i++) {
        if (i == 2 || i == 6 || i == 12) print_line(40);
        std::cout << kTotalNames[i] << " = " << v[i] << std::endl;
    }
    print_line(60, "*");
}

std::string graphlet_core::get_graphlet_counts() {
    const unsigned long long v[PGD_NTOTALS] = {
        total_2_1edge, total_2_indep, total_3_tris, total_2_star, total_3_1edge, total_3_indep,
        total_4_clique, total_4_chordcycle, total_4_tailed_tris, total_4_cycle, total_3_star,
        total_4_path, total_4_1edge, total_4_2edge, total_4_2star, total_4_tri, total_4_indep};
    std::ostringstream os;
    for (int i = 0; i < PGD_NTOTALS; i++) os << kTotalNames[i] << " = " << v[i] << "\n";
    return os.str();
}

void graphlet_core::write_macro_stats(std::string filename) {
    std::ofstream f(filename.c_str());
    f << get_graphlet_counts();
}

std::string graphlet_core::get_micro_stats(bool output_id, std::string delim) {
    std::string out;
    if (output_id) out += "src" + delim + "dst" + delim;
    out += "triangle" + delim + "2-star" + delim + "4-clique" + delim + "4-chordal-cycle" + delim +
           "4-tailed-triangle" + delim + "4-cycle" + delim + "3-star" + delim + "4-path" + "\n";
    std::vector<char> buf(256);
    out.reserve(out.size() + (size_t)m_ * 40);
    for (size_t e = 0; e < E_ordered.size() && e < tri.size(); e++) {
        char* p = buf.data();
        const int32_t id = E_ordered[e];
        const unsigned long long vals[10] = {(unsigned long long)e_v[(size_t)id] + 1,
                                             (unsigned long long)e_u[(size_t)id] + 1,
                                             tri[e], two_star[e], local_4_clique[e],
                                             local_4_chordal_cycle[e], local_4_tailed_tris[e],
                                             local_4_cycle[e], local_3_star[e],
                                             local_4_path[e]};
        for (int k = output_id ? 0 : 2; k < 10; k++) {
            p = put_u64(p, vals[k]);
            if (k < 9) for (char c : delim) *p++ = c;
        }
        *p++ = '\n';
        out.append(buf.data(), (size_t)(p - buf.data()));
    }
    return out;
}

void graphlet_core::print_micro_stats(bool output_id, std::string delim) {
    print_line(80);
    std::cout << "\n\nMicro graphlet statistics" << std::endl;
    print_line(80);
    if (delim == ",") {
        // formatted on the device and written straight to stdout
        std::cout.flush();
        std::fflush(stdout);
        check(pgd_write_micro(g_, 1, output_id ? 1 : 0, "", nullptr), "print_micro_stats");
        std::cout << std::endl;
    } else {
        std::cout << get_micro_stats(output_id, delim) << std::endl;
    }
}

void graphlet_core::write_micro_stats(std::string& filename, bool output_id) {
    const int fd = ::open(filename.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) return;  // as an ofstream that failed to open: nothing is written
    const pgd_status st = pgd_write_micro(g_, fd, output_id ? 1 : 0, "% ", nullptr);
    ::close(fd);
    check(st, "write_micro_stats");
}

namespace {
std::string gfd_block(const std::vector<unsigned long long>& G4, const char* const* names, bool guard_zero,
                      bool last_is_rest, std::vector<double>* out) {
    std::ostringstream os;
    unsigned long long sum = 0;
    for (unsigned long long x : G4) sum += x;
    double cum = 0;
    std::vector<double> g(G4.size(), 0);
    for (size_t i = 0; i < G4.size(); i++) {
        if (!guard_zero || sum > 0) g[i] = (double)G4[i] / (double)sum;
        else g[i] = 0;
        cum += g[i];
        os << names[i] << "\t" << g[i] << "\n";
    }
    if (last_is_rest) g[G4.size() - 1] = 1.0 - cum;
    os << "\n";
    if (out) *out = g;
    return os.str();
}
}  // namespace

// GFD (graphlet_core.cpp:1955-2056): display names as the reference prints them
std::string graphlet_core::compute_GFD() {
    static const char* names[] = {"4-clique     ",   "4-chordal-cycle", "4-tailed-tri", "4-cycle       ",
                                  "3-star       ",   "4-path       ",   "4-node-1-tri   ", "4-node-2-star  ",
                                  "4-node-2-edge",   "4-node-1-edge  ", "4-node-indep  "};
    std::vector<unsigned long long> G4 = {total_4_clique, total_4_chordcycle, total_4_tailed_tris, total_4_cycle,
                                          total_3_star,   total_4_path,       total_4_tri,         total_4_2star,
                                          total_4_2edge,  total_4_1edge,      total_4_indep};
    return gfd_block(G4, names, false, true, nullptr);
}

std::string graphlet_core::compute_connected_GFD() {
    static const char* names[] = {"4-clique     ", "4-chordal-cycle", "4-tailed-tri", "4-cycle       ",
                                  "3-star       ", "4-path       "};
    std::vector<unsigned long long> G4 = {total_4_clique, total_4_chordcycle, total_4_tailed_tris,
                                          total_4_cycle,  total_3_star,       total_4_path};
    return gfd_block(G4, names, true, false, nullptr);
}

std::string graphlet_core::compute_disconnected_GFD() {
    static const char* names[] = {"4-node-1-tri   ", "4-node-2-star", "4-node-2-edge  ", "4-node-1-edge  ",
                                  "4-node-indep  "};
    std::vector<unsigned long long> G4 = {total_4_tri, total_4_2star, total_4_2edge, total_4_1edge, total_4_indep};
    return gfd_block(G4, names, true, true, nullptr);
}

void graphlet_core::print_GFD() {
    print_line(80);
    std::cout << "Graphlet Frequency Distribution (GFD)\n";
    print_line(80);
    std::cout << compute_GFD() << std::endl;
}

void graphlet_core::print_connected_GFD() {
    print_line(80);
    std::cout << "Connected Motif Graphlet Frequency Distribution (GFD)\n";
    print_line(80);
    std::cout << compute_connected_GFD() << std::endl;
}

void graphlet_core::print_disconnected_GFD() {
    print_line(80);
    std::cout << "Disconnected Motif Graphlet Frequency Distribution (GFD)\n";
    print_line(80);
    std::cout << compute_disconnected_GFD() << std::endl;
}

}  // namespace pgd_amd
