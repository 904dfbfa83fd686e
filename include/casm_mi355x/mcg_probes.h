/* mcg_probes.h -- measured bandwidth peaks (casmcode_monte_amd/lib/libcasm_monte_probes.so).
 *
 * Microbenchmarks with no counterpart in the reference: they measure the
 * denominators of the roofline figures the benchmark reports (HBM streaming
 * bandwidth and aggregate LDS read bandwidth of the current device) beside
 * the datasheet peaks.  They run on the current device and the null stream,
 * return 0 on success and 1 on failure (mcgp_last_error() has the message).
 */
#ifndef CASM_MI355X_MCG_PROBES_H
#define CASM_MI355X_MCG_PROBES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Copy `bytes` (>= 1 MiB, a multiple of 16) between two device buffers `reps`
 * times with 16-byte loads and stores; *gbs = 2 * bytes / time per copy. */
int mcgp_hbm_copy(int64_t bytes, int64_t reps, double *gbs, double *ms_per_copy);

/* Two 1024-thread workgroups per CU read 16-byte LDS words without bank
 * conflicts, 8 per thread per iteration; *gbs = bytes read / time. */
int mcgp_lds_read(int64_t iters, double *gbs, double *ms);

const char *mcgp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CASM_MI355X_MCG_PROBES_H */
