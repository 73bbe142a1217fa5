// pgd_batch.hip -- batched and served forms (placeholder until written).
#include "pgd_internal.hpp"
extern "C" {
pgd_status pgd_batch_create(pgd_ctx*, int32_t, const int32_t*, const int64_t*, const int32_t*, const int32_t*, uint32_t, pgd_batch**) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_run(pgd_batch*, uint64_t*, uint64_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_micro(pgd_batch*, int32_t, int32_t, uint64_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_info(pgd_batch*, int32_t*, int64_t*, int64_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_free(pgd_batch*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_set_tuning(pgd_batch*, const pgd_tuning*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_get_tuning(pgd_batch*, pgd_tuning*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_autotune(pgd_batch*, int32_t, pgd_tuning*, double*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_batch_last_ms(pgd_batch*, double*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_create(pgd_ctx*, int64_t, int32_t, pgd_server**) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_submit(pgd_server*, int32_t, int64_t, const int32_t*, const int32_t*, uint32_t, int64_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_poll(pgd_server*, int64_t*, int32_t, int32_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_result(pgd_server*, int64_t, uint64_t*, uint64_t*, int64_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_micro(pgd_server*, int64_t, int32_t, uint64_t*) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_release(pgd_server*, int64_t) { return PGD_ERR_UNSUPPORTED; }
pgd_status pgd_server_free(pgd_server*) { return PGD_ERR_UNSUPPORTED; }
}
