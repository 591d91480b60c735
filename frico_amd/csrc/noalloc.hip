// Field evaluation that allocates nothing (frico_fields_noalloc).
//
// The triangle currents the field sums read (4 complex128 values per triangle and current vector,
// the RWG -> triangle scatter of src/emw_solver.h:137-164) live in `scratch`, a file-scope device
// array of FRICO_FIELDS_NOALLOC_MAX_BYTES.  Each GPU the library runs on holds its own copy at its
// own address until the process ends.  In host code `scratch` only names the array: its device
// address is resolved once per device with hipGetSymbolAddress (noalloc_init, called by
// frico_ctx_create with the device current), and only that address reaches kernels.
//
// Every call on one GPU shares that GPU's array: calls from contexts whose streams overlap on one
// device race on it.  Calls on one context are ordered by its stream.  The scatter writes every
// value the field sums then read, so nothing has to start at zero and the last call's contents
// are harmless.
#include "ctx.hpp"

#include <mutex>

static __device__ __align__(256) unsigned char scratch[FRICO_FIELDS_NOALLOC_MAX_BYTES];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr; // per device: the device address of `scratch` (null until resolved)

void* scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}
} // namespace

namespace frico {

// Resolves the device address of `scratch` on `device`, which must be current.  Once per device.
int noalloc_init(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0)
        return set_error("bad device %d", device), -1;
    if ((int)g_scratch_ptr.size() <= device)
        g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[device])
        return 0;
    void* p = nullptr;
    FRICO_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p)
        return set_error("no device address for the field scratch on device %d", device), FRICO_ERR_HIP;
    g_scratch_ptr[device] = p;
    return 0;
}

} // namespace frico

extern "C" {

int frico_fields_noalloc(frico_ctx* c, double freq, const double* currents, int64_t ldi, int64_t ncols,
                         const double* pts, int64_t npts, const int32_t* cols, double h_sign, double* E, double* H)
{
    if (!c)
        return set_error("null context"), -1;
    if (!c->have_mesh)
        return set_error("no mesh loaded"), FRICO_ERR_STATE;
    if (!(freq > 0))
        return set_error("frequency must be positive"), -2;
    if (!currents)
        return set_error("null current vectors"), -3;
    if (ldi < c->N)
        return set_error("leading dimension %lld below N = %lld", (long long)ldi, (long long)c->N), -4;
    if (ncols < 1 || ncols > 65535)
        return set_error("ncols = %lld outside 1..65535", (long long)ncols), -5;
    const int64_t T = c->hm.nt;
    const double need = 64.0 * (double)T * (double)ncols; // 4 complex128 values per triangle and column
    if (need > (double)FRICO_FIELDS_NOALLOC_MAX_BYTES)
        return set_error("triangle currents of %lld triangles x %lld columns need %.0f bytes, above the %lld-byte "
                         "scratch of frico_fields_noalloc (pass fewer columns per call)",
                         (long long)T, (long long)ncols, need, (long long)FRICO_FIELDS_NOALLOC_MAX_BYTES),
               FRICO_ERR_LIMIT;
    if (!pts)
        return set_error("null points"), -6;
    if (npts < 1 || npts > (1 << 30))
        return set_error("npts = %lld outside 1..2^30", (long long)npts), -7;
    if (!E && !H)
        return set_error("null outputs"), -10;
    void* ts = scratch_ptr(c->device);
    if (!ts)
        return set_error("field scratch not resolved on device %d", c->device), FRICO_ERR_STATE;
    FRICO_HIP_CHECK(hipSetDevice(c->device));
    const double omega = 2.0 * PI * freq;
    const double k = omega * std::sqrt(MU0 * EPS0);
    int rc = launch_scatter(c->dm, reinterpret_cast<const c128*>(currents), ldi, (int)ncols, static_cast<c128*>(ts),
                            c->st);
    if (rc)
        return rc;
    return launch_fields(c->dm, k, omega * MU0, static_cast<const c128*>(ts), (int)ncols, pts, cols, (int)npts,
                         h_sign >= 0 ? 1.0 : -1.0, reinterpret_cast<c128*>(E), reinterpret_cast<c128*>(H), c->st);
}

} // extern "C"
