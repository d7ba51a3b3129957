// rd_bench.hip -- measurement support: timed graph of K steps, copy kernel.
#include "../../include/rdgpu_bench.h"
#include "rd_host.h"
#include "rd_kernels.h"

#include <chrono>
#include <string>
#include <vector>


__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    size_t stride = (size_t)gridDim.x * 256;
    for (; i < n; i += stride) b[i] = a[i];
}

#define BCK(x)                                                              \
    do {                                                                    \
        hipError_t e_ = (x);                                                \
        if (e_ != hipSuccess) {                                             \
            h->err = std::string(#x) + ": " + hipGetErrorString(e_);        \
            rc = RD_EHIP;                                                   \
            goto done;                                                      \
        }                                                                   \
    } while (0)

extern "C" int rd_bench_steps(rd_system* h, int nsteps, double* out)
{
    if (!h || nsteps < 1 || !out) return RD_EINVAL;
    int rc = rd_host_refresh(h);
    if (rc) return rc;
    std::vector<hipEvent_t> ev(5 * (size_t)nsteps, nullptr);
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    double tf = 0.0, ts = 0.0, te = 0.0;
    const bool ext = h->ext_active();
    float ms = 0.0f;
    std::chrono::steady_clock::time_point t0, t1;
    for (auto& e : ev) BCK(hipEventCreate(&e));
    // pass 1: the K steps as one graph, host wall clock around launch + sync
    BCK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < nsteps; ++i) {
        if (ext) rd_host_launch_ext(h);
        rd_host_launch_force(h);
        rd_host_launch_sort(h, 1);
    }
    BCK(hipStreamEndCapture(h->stream, &g));
    BCK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    BCK(hipDeviceSynchronize());
    t0 = std::chrono::steady_clock::now();
    BCK(hipGraphLaunch(ge, h->stream));
    BCK(hipStreamSynchronize(h->stream));
    t1 = std::chrono::steady_clock::now();
    out[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    h->step += (uint64_t)nsteps;
    // pass 2: K more steps launched directly with events between the kernel
    // groups (events recorded inside a graph have no timestamps in HIP); the
    // force kernel's own events are timestamped by its dispatch
    // (hipExtLaunchKernelGGL), so their interval is the kernel's duration
    for (int i = 0; i < nsteps; ++i) {
        BCK(hipEventRecord(ev[5 * i], h->stream));
        if (ext) rd_host_launch_ext(h);
        BCK(hipEventRecord(ev[5 * i + 1], h->stream));
        rd_host_launch_force(h, ev[5 * i + 2], ev[5 * i + 3]);
        rd_host_launch_sort(h, 1);
        BCK(hipEventRecord(ev[5 * i + 4], h->stream));
    }
    BCK(hipStreamSynchronize(h->stream));
    h->step += (uint64_t)nsteps;
    for (int i = 0; i < nsteps; ++i) {
        BCK(hipEventElapsedTime(&ms, ev[5 * i], ev[5 * i + 1]));
        te += ms;
        BCK(hipEventElapsedTime(&ms, ev[5 * i + 2], ev[5 * i + 3]));
        tf += ms;
        BCK(hipEventElapsedTime(&ms, ev[5 * i + 3], ev[5 * i + 4]));
        ts += ms;
    }
    BCK(hipEventElapsedTime(&ms, ev[0], ev[5 * (size_t)nsteps - 1]));
    out[1] = ms;
    out[2] = tf;
    out[3] = ts;
    out[4] = te;
done:
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

extern "C" int rd_bench_copy(rd_system* h, long bytes, int reps, double* avg_ms)
{
    if (!h || bytes < 16 || reps < 1) return RD_EINVAL;
    int rc = RD_OK;
    size_t n = (size_t)bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    BCK(hipSetDevice(h->dev));
    BCK(hipMalloc(&a, n * 16));
    BCK(hipMalloc(&b, n * 16));
    BCK(hipMemsetAsync(a, 1, n * 16, h->stream));
    BCK(hipEventCreate(&e0));
    BCK(hipEventCreate(&e1));
    BCK(hipEventRecord(e0, h->stream));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, h->stream, a, b, n);
    BCK(hipEventRecord(e1, h->stream));
    BCK(hipEventSynchronize(e1));
    BCK(hipEventElapsedTime(&ms, e0, e1));
    if (avg_ms) *avg_ms = ms / reps;
done:
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int rd_bench_ext(rd_system* h, int reps, double* avg_ms)
{
    if (!h || reps < 1 || !avg_ms) return RD_EINVAL;
    int rc = rd_host_refresh(h);
    if (rc) return rc;
    if (!h->ext_active()) {
        h->err = "no external pass (set a composite first)";
        return RD_ESTATE;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms = 0.0f;
    BCK(hipEventCreate(&e0));
    BCK(hipEventCreate(&e1));
    rd_host_launch_ext(h);                   // warm-up
    BCK(hipEventRecord(e0, h->stream));
    for (int r = 0; r < reps; ++r) rd_host_launch_ext(h);
    BCK(hipEventRecord(e1, h->stream));
    BCK(hipEventSynchronize(e1));
    BCK(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = ms / reps;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

extern "C" int rd_bench_steps_direct(rd_system* h, int nsteps, double* out)
{
    if (!h || nsteps < 1 || !out) return RD_EINVAL;
    int rc = rd_host_refresh(h);
    if (rc) return rc;
    std::vector<hipEvent_t> ev(5 * (size_t)nsteps, nullptr);
    double tf = 0.0, ts = 0.0, te = 0.0;
    const bool ext = h->ext_active();
    float ms = 0.0f;
    std::chrono::steady_clock::time_point t0, t1;
    for (auto& e : ev) BCK(hipEventCreate(&e));
    BCK(hipStreamSynchronize(h->stream));
    t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < nsteps; ++i) {
        BCK(hipEventRecord(ev[5 * i], h->stream));
        if (ext) rd_host_launch_ext(h);
        BCK(hipEventRecord(ev[5 * i + 1], h->stream));
        rd_host_launch_force(h, ev[5 * i + 2], ev[5 * i + 3]);
        rd_host_launch_sort(h, 1);
        BCK(hipEventRecord(ev[5 * i + 4], h->stream));
    }
    BCK(hipStreamSynchronize(h->stream));
    t1 = std::chrono::steady_clock::now();
    h->step += (uint64_t)nsteps;
    for (int i = 0; i < nsteps; ++i) {
        BCK(hipEventElapsedTime(&ms, ev[5 * i], ev[5 * i + 1]));
        te += ms;
        BCK(hipEventElapsedTime(&ms, ev[5 * i + 2], ev[5 * i + 3]));
        tf += ms;
        BCK(hipEventElapsedTime(&ms, ev[5 * i + 3], ev[5 * i + 4]));
        ts += ms;
    }
    out[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    BCK(hipEventElapsedTime(&ms, ev[0], ev[5 * (size_t)nsteps - 1]));
    out[1] = ms;
    out[2] = tf;
    out[3] = ts;
    out[4] = te;
done:
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

