/*
 * gsph_probe.h - measurement-only library (libgsph_probe.so), separate from
 * the product library libgsph.so.  No counterpart in the reference.  The two
 * kernels calibrate the roofline figures bench.py reports for the SPH step:
 * a copy of known size (its HBM byte counters scale FETCH_SIZE / WRITE_SIZE)
 * and a VALU peak measurement (independent FP32 FMA chains).  Status codes
 * are those of gsph.h; gsph_probe_last_error() returns the last message.
 */
#ifndef GSPH_PROBE_H_
#define GSPH_PROBE_H_

#ifdef __cplusplus
extern "C" {
#endif

const char *gsph_probe_last_error(void);
/* A plain float4 copy kernel over `bytes` bytes, run reps (<= 1000) times;
 * *ms receives the average duration of one launch. */
int gsph_probe_copy(double bytes, int reps, float *ms);
/* Eight independent FP32 FMA chains per lane on 8 waves per SIMD, iters x 64
 * FMAs per lane (iters <= 2^20, about 1 s).  Returns the launch's ms and its
 * lane operations (threads x FMAs): the peak VALU rate against which the
 * density and force kernels' active-lane operations are measured. */
int gsph_probe_valu(int iters, float *ms, double *lane_ops);

#ifdef __cplusplus
}
#endif

#endif /* GSPH_PROBE_H_ */
