// pgd_engine_orient.hip -- degree-oriented engine (placeholder until written).
#include "pgd_internal.hpp"
namespace pgd {
struct OrientState {};
void engine_orient(pgd_ctx*, const GraphView&, EdgeCounts, Timing*, OrientState**) {
    throw Error{PGD_ERR_UNSUPPORTED, "oriented engine not built"};
}
void orient_state_free(pgd_ctx*, OrientState* s) { delete s; }
}  // namespace pgd
