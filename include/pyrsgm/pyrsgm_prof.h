/*
 * pyrsgm_prof.h -- stage timing hook (measurement only).
 *
 * No counterpart in gegeges/pyrsgm.  While enabled, every per-level launch
 * sequence records a HIP event pair around each of its kernels on the stream
 * it runs on; the bench reads the pairs back to get per-kernel durations in
 * the same run it times.  Disabled (the default) it costs one branch per
 * launch.  Not thread-safe: enable it only around single-threaded benchmarks.
 */
#ifndef PYRSGM_PROF_H
#define PYRSGM_PROF_H

#ifdef __cplusplus
extern "C" {
#endif

/* Stage ids reported by pyrsgm_prof_get. */
#define PYRSGM_STAGE_PREP     0  /* census words / BT intervals + key reset */
#define PYRSGM_STAGE_COST     1  /* matching-cost volume                     */
#define PYRSGM_STAGE_PATHS    2  /* L->R, R->L and T->B paths (one launch)   */
#define PYRSGM_STAGE_WTA      3  /* path sums + winner-take-all              */
#define PYRSGM_STAGE_POST     4  /* LR check + median + outputs              */
#define PYRSGM_STAGE_PYRDOWN  5  /* pyramid down-sampling (both views)       */

/* Enable recording of up to max_records event pairs (pre-created here). */
int pyrsgm_prof_start(int max_records);
/* Record only launches of this stage at this level width/height (-1: any).
 * Persists across start/stop; (-1, -1, -1) records everything. */
void pyrsgm_prof_filter(int stage, int width, int height);
/* Stop recording; recorded pairs stay readable until the next start. */
void pyrsgm_prof_stop(void);
int pyrsgm_prof_count(void);
/* Waits for record i and returns its stage, level geometry and duration. */
int pyrsgm_prof_get(int i, int* stage, int* width, int* height, int* labels, float* ms);

#ifdef __cplusplus
}
#endif
#endif
