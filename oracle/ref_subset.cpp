// ref_subset.cpp -- runs the reference library's sampled decomposition,
// for golden vectors only (built by `make -C oracle ref` into oracle/_ref/).
//
// The reference driver (graphlet_driver.cpp:53-92) always orders all edges;
// graphlet_core also has sort_edges(ordering, s2l, E_s) for a sampled edge
// set (graphlet_core.cpp:836-851), reachable only through the library.  This
// program follows the driver's sequence with that overload:
//
//   ref_subset SUBSET_FILE <pgd flags: -f G -o ORD [--s2l] -c X.macro -m X.micro ...>
//
// SUBSET_FILE lists edge ids (whitespace separated).  Written in C++98 to
// build with the reference's own flags.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "graphlet_core.h"

using namespace graphlet;

int main(int argc, char* argv[]) {
    if (argc < 3) {
        std::cerr << "usage: ref_subset SUBSET_FILE <pgd flags>" << std::endl;
        return 2;
    }
    std::vector<unsigned long> E_s;
    {
        std::ifstream f(argv[1]);
        if (!f) {
            std::cerr << "cannot read " << argv[1] << std::endl;
            return 2;
        }
        unsigned long x;
        while (f >> x) E_s.push_back(x);
    }
    // params reads argv[1..]: the subset file takes the program name's place
    params p(argc - 1, argv + 1);
    graphlet_core G(p);
    G.compute_cores();
    for (size_t i = 0; i < E_s.size(); i++)
        if (E_s[i] >= (unsigned long)G.num_edges()) {
            std::cerr << "edge id out of range" << std::endl;
            return 2;
        }
    G.sort_edges(p.ordering, p.is_small_to_large, E_s);
    if (p.is_micro_stats()) G.graphlet_decomposition_micro(p.workers);
    else G.graphlet_decomposition(p.workers);
    if (p.is_macro_stats()) G.write_macro_stats(p.macro_stats_filename);
    if (p.is_micro_stats()) G.write_micro_stats(p.micro_stats_filename);
    return 0;
}
