/* hlanc_probes.h -- measurement probes of libhlanc_hip.so.
 *
 * No counterpart in the reference: these time the memory and launch
 * behaviour the kernels' design rests on (read and gather rates, CGS pass
 * time, grid-barrier and graph-node costs), on buffers of their own.  They
 * are not part of the solver interface (hlanc_capi.h) and solve nothing.
 * Return values are the hl_status codes of hlanc_capi.h. */
#ifndef HLANC_PROBES_H
#define HLANC_PROBES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Average seconds per launch of one fused CGS pass (w -= V h; h' = V^T w) over
 * an n x m FP64 basis, layout tiled (1) or column-major (0), timed with HIP
 * events over `reps` launches on the default stream (own buffers). */
int hl_bench_cgs(int dev, int n, int m, int tiled, int reps, double *seconds);
/* Average seconds per pass of a read-only 16-byte-load stream over `bytes`
 * bytes of HBM (own buffer): the read rate this device reaches. */
int hl_bench_stream(int dev, long long bytes, int reps, double *seconds);
/* Average seconds per pass of a gather probe: m 4-byte indices streamed, each
 * followed by an 8-byte read of x[index] from a vector of nx doubles; indices
 * uniformly random (random = 1) or i mod nx (0).  The SpMV's x-gather pattern
 * without the matrix values. */
int hl_bench_gather(int dev, long long nx, long long m, int random, int reps, double *seconds);
/* Seconds per grid-wide barrier of a cooperative launch of `blocks` blocks of
 * 256 threads running `barriers` barriers (release add on one agent-scope
 * counter, acquire loads until every block has arrived; between barriers each
 * block writes 256 doubles and reads its neighbour's).  Every wait is bounded:
 * HL_ERR_STATE if one expires or a block read stale data, HL_ERR_ARG if the
 * blocks cannot all be co-resident on the device (no launch then). */
int hl_bench_grid_barrier(int dev, int blocks, int barriers, double *seconds);
/* Seconds per node of a HIP graph of `nodes` empty kernels of `blocks` blocks,
 * captured from one stream and replayed `reps` times. */
int hl_bench_graph_nodes(int dev, int nodes, int blocks, int reps, double *seconds);

#ifdef __cplusplus
}
#endif
#endif
