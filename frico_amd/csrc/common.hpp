// Shared device/host definitions of the MI355X EFIE library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <tuple>
#include <utility>
#include <vector>

namespace frico {

// src/constants.h:23-24
constexpr double MU0 = 1.256637061435917e-06;
constexpr double EPS0 = 8.8541878188e-12;
constexpr double PI = 3.14159265358979323846;
constexpr double INV_PI = 0.31830988618379067154;

// Interleaved complex (re, im), the memory layout of std::complex<T>.
template <typename T> struct alignas(2 * sizeof(T)) cx {
    T re, im;
};
using c64 = cx<float>;
using c128 = cx<double>;

template <typename T> __host__ __device__ inline cx<T> mk(T r, T i) { return cx<T>{r, i}; }
template <typename T> __host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b)
{
    return {a.re + b.re, a.im + b.im};
}
template <typename T> __host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b)
{
    return {a.re - b.re, a.im - b.im};
}
template <typename T> __host__ __device__ inline cx<T> operator*(cx<T> a, cx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ inline cx<T> operator*(T s, cx<T> a) { return {s * a.re, s * a.im}; }
template <typename T> __host__ __device__ inline cx<T> conj(cx<T> a) { return {a.re, -a.im}; }
// Smith's algorithm for a / b (R. L. Smith, CACM 5(8), 1962).
template <typename T> __host__ __device__ inline cx<T> cdiv(cx<T> a, cx<T> b)
{
    T ar = b.re < 0 ? -b.re : b.re, ai = b.im < 0 ? -b.im : b.im;
    if (ar >= ai) {
        T r = b.im / b.re, d = b.re + b.im * r;
        return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
    }
    T r = b.re / b.im, d = b.im + b.re * r;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
template <typename T> __host__ __device__ inline T cabs1(cx<T> a)
{
    return (a.re < 0 ? -a.re : a.re) + (a.im < 0 ? -a.im : a.im);
}

} // namespace frico

#define FRICO_HIP_CHECK(expr)                                                                     \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return ::frico::hip_fail(_e, #expr, __FILE__, __LINE__);                              \
    } while (0)

namespace frico {
// Element strides between the systems of a batched launch, the system being blockIdx.z: a kernel
// that takes a Batch offsets its pointers by z times the stride it assigns to each (all zero for
// one system).  Each such kernel states which pointer takes which stride.
struct Batch {
    int64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
};

int hip_fail(hipError_t e, const char* what, const char* file, int line);
void set_error(const char* fmt, ...);
// Raises a kernel's dynamic-LDS limit to `bytes` (gfx950 CUs have 160 KB, shared with static LDS).
// Each (device, kernel) is set once, so that later launches, including launches under HIP graph
// capture, make no attribute call.
hipError_t allow_lds(const void* func, size_t bytes);

// Graph building without stream capture.  While a recorder is installed on the calling thread,
// every launch_kernel on the recorder's stream becomes a kernel node of recorder.graph, each node
// depending on the previous one, instead of running.  Recording is per thread and touches no
// stream state, so other threads keep launching and allocating freely.  Only kernel launches are
// recorded: a sequence recorded this way must not contain copies, memsets or event operations.
struct GraphRecorder {
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphNode_t last = nullptr;
    hipError_t err = hipSuccess;
};
GraphRecorder*& graph_recorder(); // this thread's recorder, or null

template <typename... P, typename... A, size_t... I>
inline hipError_t launch_kernel_impl(void (*k)(P...), dim3 grid, dim3 block, size_t sh, hipStream_t st,
                                     std::index_sequence<I...>, A&&... a)
{
    std::tuple<std::decay_t<P>...> vals(static_cast<std::decay_t<P>>(std::forward<A>(a))...);
    void* argv[sizeof...(P) > 0 ? sizeof...(P) : 1] = {(void*)&std::get<I>(vals)...};
    GraphRecorder* r = graph_recorder();
    if (!r || r->st != st)
        return hipLaunchKernel(reinterpret_cast<const void*>(k), grid, block, argv, sh, st);
    hipKernelNodeParams p{};
    p.func = reinterpret_cast<void*>(k);
    p.gridDim = grid;
    p.blockDim = block;
    p.sharedMemBytes = (unsigned)sh;
    p.kernelParams = argv;
    p.extra = nullptr;
    hipGraphNode_t node = nullptr;
    const hipError_t e = hipGraphAddKernelNode(&node, r->graph, r->last ? &r->last : nullptr, r->last ? 1 : 0, &p);
    if (e == hipSuccess)
        r->last = node;
    else if (r->err == hipSuccess)
        r->err = e;
    return e;
}

// Launches kernel k on st, or appends it to this thread's graph recorder for st.  A failed launch
// is reported by the next hipGetLastError like hipLaunchKernelGGL's.
template <typename... P, typename... A>
inline void launch_kernel(void (*k)(P...), dim3 grid, dim3 block, size_t sh, hipStream_t st, A&&... a)
{
    static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
    (void)launch_kernel_impl(k, grid, block, sh, st, std::index_sequence_for<P...>{}, std::forward<A>(a)...);
}

} // namespace frico

namespace frico {

// Brackets kernel launches with HIP events on their stream and accumulates algorithmic work.
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> ev;
    std::vector<double> work, bytes;
    size_t used = 0;
    double total_ms = 0, total_work = 0, total_bytes = 0;
    long launches = 0;

    void begin(hipStream_t st)
    {
        if (!enabled)
            return;
        if (used + 2 > ev.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
                return;
            ev.push_back(a);
            ev.push_back(b);
        }
        (void)hipEventRecord(ev[used], st);
    }
    // w: algorithmic work of the bracketed launches (flops); b: their algorithmic device bytes
    void end(hipStream_t st, double w, double b = 0.0)
    {
        if (!enabled || used + 2 > ev.size())
            return;
        (void)hipEventRecord(ev[used + 1], st);
        work.resize(used / 2 + 1);
        bytes.resize(used / 2 + 1);
        work[used / 2] = w;
        bytes[used / 2] = b;
        used += 2;
    }
    // Waits for the recorded events and folds them into the totals.
    void collect()
    {
        for (size_t i = 0; i + 1 < used; i += 2) {
            float ms = 0;
            (void)hipEventSynchronize(ev[i + 1]);
            (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
            total_ms += ms;
            total_work += work[i / 2];
            total_bytes += bytes[i / 2];
            launches++;
        }
        used = 0;
    }
    void reset()
    {
        collect();
        total_ms = total_work = total_bytes = 0;
        launches = 0;
    }
    ~KernelTimer()
    {
        for (auto e : ev)
            (void)hipEventDestroy(e);
    }
};

} // namespace frico
