// pgd_engine_dense.hip -- dense-tier engine (placeholder until written).
#include "pgd_internal.hpp"
namespace pgd {
bool engine_dense_ok(const GraphView&) { return false; }
void engine_dense(pgd_ctx*, const GraphView&, EdgeCounts, Timing*) {
    throw Error{PGD_ERR_UNSUPPORTED, "dense engine not built"};
}
}  // namespace pgd
