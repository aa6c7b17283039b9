// Batched device-resident chains (persistent kernel filled in next).
#include <hip/hip_runtime.h>
#include "fastmc_amd.h"
extern "C" {
int fmc_batch_create(fmc_batch**, int, int) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_destroy(fmc_batch*) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_add_chain(fmc_batch*, const fmc_ctx*, const fmc_chain_params*, int*) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_remove_chain(fmc_batch*, int) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_run(fmc_batch*, long long) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_launch(fmc_batch*, long long) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_sync(fmc_batch*) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_stats(fmc_batch*, int, fmc_chain_stats*) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_enable_log(fmc_batch*, int, long long) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_get_log(fmc_batch*, int, long long, double*) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_resync(fmc_batch*, double*) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_set_tile(fmc_batch*, int) { return FMC_ERR_UNSUPPORTED; }
int fmc_batch_get_tile(const fmc_batch*, int*) { return FMC_ERR_UNSUPPORTED; }
int fmc_chain_footprint(const fmc_ctx*, int, long long*) { return FMC_ERR_UNSUPPORTED; }
}
