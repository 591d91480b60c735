// Shared device/host definitions of the MI355X EFIE library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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
int hip_fail(hipError_t e, const char* what, const char* file, int line);
void set_error(const char* fmt, ...);
// Raises a kernel's dynamic-LDS limit to `bytes` (gfx950 CUs have 160 KB, shared with static LDS).
inline hipError_t allow_lds(const void* func, size_t bytes)
{
    return hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
} // namespace frico

namespace frico {

// Brackets kernel launches with HIP events on their stream and accumulates algorithmic work.
struct KernelTimer {
    bool enabled = false;
    std::vector<hipEvent_t> ev;
    std::vector<double> work;
    size_t used = 0;
    double total_ms = 0, total_work = 0;
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
    void end(hipStream_t st, double w)
    {
        if (!enabled || used + 2 > ev.size())
            return;
        (void)hipEventRecord(ev[used + 1], st);
        work.resize(used / 2 + 1);
        work[used / 2] = w;
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
            launches++;
        }
        used = 0;
    }
    void reset()
    {
        collect();
        total_ms = total_work = 0;
        launches = 0;
    }
    ~KernelTimer()
    {
        for (auto e : ev)
            (void)hipEventDestroy(e);
    }
};

} // namespace frico
