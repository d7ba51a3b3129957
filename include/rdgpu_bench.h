/*
 * rdgpu_bench.h -- measurement entry points of libCReaDDyGPU.so.
 * These have no counterpart in the reference: they exist for the benchmark
 * and for profiler calibration.
 */
#ifndef RDGPU_BENCH_H
#define RDGPU_BENCH_H
#include "rdgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run nsteps steps as one captured graph (timed on the host), then nsteps
 * more steps launched directly with HIP events between the kernel groups of
 * every step (graph nodes carry no event timestamps); the force/update
 * kernel's events are timestamped by its own dispatch (hipExtLaunchKernelGGL).
 * out[0] = host wall time of graph launch + synchronize (ms),
 * out[1] = first event to last of pass 2 (ms), out[2] = sum of the force/update
 * kernel durations, out[3] = sum of cell-sort times (force kernel end to the
 * event after the scatter), out[4] = sum of external-pass times (composite or
 * surrogate; 0 without).  out holds 5 values. */
int rd_bench_steps(rd_system* h, int nsteps, double* out);

/* The same measurement without graph capture: nsteps steps launched directly
 * with the events of pass 2 of rd_bench_steps; out[0] = host wall time of the
 * launches + synchronize, out[1..4] as rd_bench_steps.  For profilers that
 * should not go through graph capture and launch. */
int rd_bench_steps_direct(rd_system* h, int nsteps, double* out);

/* Copy `bytes` bytes device-to-device with a plain vector-copy kernel, reps
 * times, on the handle's stream (profiler calibration of HBM counters).
 * ms: average kernel time. */
int rd_bench_copy(rd_system* h, long bytes, int reps, double* ms);

/* The external pass alone (exact composite or surrogate, whichever the step
 * uses) on the current state, reps launches; ms: average.  The positions do
 * not move, so this times a pass that the step itself could not run, e.g. a
 * composite whose step diverges. */
int rd_bench_ext(rd_system* h, int reps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
