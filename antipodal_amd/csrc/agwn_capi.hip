// C ABI of libagwn.so (declared in include/agwn.h): argument checks, host<->device staging, error reporting.
// No C++ exception crosses this layer; every entry point returns an agwn status.

#include "agwn_internal.hpp"
#include "../../include/agwn.h"

#include <algorithm>
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

namespace agwn
{
namespace
{
thread_local std::string g_last_error;
}

void set_error(char const* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

int ensure_device()
{
    static std::once_flag once;
    static int cached = err_nodevice;
    static std::string msg;
    std::call_once(once,
                   []
                   {
                       int n = 0;
                       hipError_t const e = hipGetDeviceCount(&n);
                       if (e != hipSuccess)
                           msg = std::string("no HIP device: ") + hipGetErrorString(e);
                       else if (n <= 0)
                           msg = "no HIP device visible";
                       else
                           cached = ok;
                   });
    if (cached != ok)
        set_error("%s", msg.c_str());
    return cached;
}

// A per-call stream for the synchronous host-pointer entry points.
struct call_stream
{
    hipStream_t s = nullptr;
    ~call_stream()
    {
        if (s)
            (void)hipStreamDestroy(s);
    }
    int create()
    {
        AGWN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        return ok;
    }
};

template <class F>
int guarded(F&& f)
{
    clear_error();
    try
    {
        return f();
    }
    catch (std::bad_alloc const&)
    {
        set_error("host allocation failed");
        return err_alloc;
    }
    catch (std::exception const& e)
    {
        set_error("%s", e.what());
        return err_invalid;
    }
    catch (...)
    {
        set_error("unknown C++ exception");
        return err_invalid;
    }
}

template <class T>
int frac_host(T const* pos, int64_t n, T const* seg7, int64_t nseg, T const* x0, T* out)
{
    AGWN_CHECK(n >= 0 && nseg >= 0, "negative count (n=%lld, nseg=%lld)", (long long)n, (long long)nseg);
    AGWN_CHECK(x0 != nullptr, "x0 is null");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    if (n == 0)
        return ok;
    AGWN_TRY(ensure_device());
    call_stream cs;
    AGWN_TRY(cs.create());
    dbuf dp, ds, dout;
    AGWN_TRY(dp.alloc(size_t(n) * 3 * sizeof(T), cs.s));
    AGWN_TRY(ds.alloc(size_t(std::max<int64_t>(nseg, 1)) * 7 * sizeof(T), cs.s));
    AGWN_TRY(dout.alloc(size_t(n) * sizeof(T), cs.s));
    AGWN_HIP(hipMemcpyAsync(dp.p, pos, size_t(n) * 3 * sizeof(T), hipMemcpyHostToDevice, cs.s));
    if (nseg > 0)
        AGWN_HIP(hipMemcpyAsync(ds.p, seg7, size_t(nseg) * 7 * sizeof(T), hipMemcpyHostToDevice, cs.s));
    double const x[3] = {double(x0[0]), double(x0[1]), double(x0[2])};
    if constexpr (sizeof(T) == 4)
        AGWN_TRY(frac_f32(dp.as<float>(), n, ds.as<float>(), nseg, x, nullptr, dout.as<float>(), cs.s));
    else
        AGWN_TRY(frac_f64(dp.as<double>(), n, ds.as<double>(), nseg, x, nullptr, dout.as<double>(), cs.s));
    AGWN_HIP(hipMemcpyAsync(out, dout.p, size_t(n) * sizeof(T), hipMemcpyDeviceToHost, cs.s));
    AGWN_HIP(hipStreamSynchronize(cs.s));
    return ok;
}

template <class T>
int frac_dev(T const* pos, int64_t n, T const* seg7, int64_t nseg, T const* x0, T* out, hipStream_t s)
{
    AGWN_CHECK(n >= 0 && nseg >= 0, "negative count (n=%lld, nseg=%lld)", (long long)n, (long long)nseg);
    AGWN_CHECK(x0 != nullptr, "x0 is null");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    if (n == 0)
        return ok;
    AGWN_TRY(ensure_device());
    double const x[3] = {double(x0[0]), double(x0[1]), double(x0[2])};
    if constexpr (sizeof(T) == 4)
        return frac_f32(pos, n, seg7, nseg, x, nullptr, out, s);
    else
        return frac_f64(pos, n, seg7, nseg, x, nullptr, out, s);
}

// math/common.hh:123-174 restated: net directed count per undirected edge, zero-sum edges dropped, oriented so the
// weight is positive. Sort-based (deterministic output order: by (lo, hi)).
template <class T>
int boundary_host(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T* out7, int64_t cap, int64_t* count)
{
    AGWN_CHECK(nv >= 0 && nidx >= 0 && cap >= 0, "negative count");
    AGWN_CHECK(nidx % 3 == 0, "index count %lld is not a multiple of 3", (long long)nidx);
    AGWN_CHECK(count != nullptr, "count is null");
    AGWN_CHECK(nidx == 0 || (verts && idx), "null mesh arrays");
    AGWN_CHECK(cap == 0 || out7, "out7 is null with cap=%lld", (long long)cap);
    std::vector<uint64_t> key(static_cast<size_t>(nidx));
    for (int64_t t = 0; t < nidx / 3; ++t)
        for (int e = 0; e < 3; ++e)
        {
            int32_t const u = idx[3 * t + e], v = idx[3 * t + (e + 1) % 3];
            AGWN_CHECK(u >= 0 && u < nv && v >= 0 && v < nv, "index out of range at triangle %lld", (long long)t);
            uint32_t const lo = uint32_t(std::min(u, v)), hi = uint32_t(std::max(u, v));
            // key: (lo, hi) in the high 62 bits, direction in bit 0 (1 = lo->hi)
            key[size_t(3 * t + e)] = ((uint64_t(lo) << 32 | hi) << 1) | (u < v ? 1u : 0u);
        }
    std::sort(key.begin(), key.end());
    int64_t n = 0;
    for (size_t i = 0; i < key.size();)
    {
        uint64_t const e = key[i] >> 1;
        int64_t net = 0;
        size_t j = i;
        for (; j < key.size() && (key[j] >> 1) == e; ++j)
            net += (key[j] & 1u) ? 1 : -1;
        if (net != 0)
        {
            if (n < cap)
            {
                int64_t const lo = int64_t(e >> 32), hi = int64_t(e & 0xffffffffu);
                int64_t const a = net > 0 ? lo : hi, b = net > 0 ? hi : lo;
                T* o = out7 + 7 * n;
                for (int k = 0; k < 3; ++k)
                {
                    o[k] = verts[3 * a + k];
                    o[3 + k] = verts[3 * b + k];
                }
                o[6] = T(net > 0 ? net : -net);
            }
            ++n;
        }
        i = j;
    }
    *count = n;
    return ok;
}
} // namespace agwn

using namespace agwn;

extern "C" {

char const* agwn_last_error(void) { return g_last_error.c_str(); }

int agwn_version(void) { return 100; }

int agwn_device_available(void)
{
    clear_error();
    return ensure_device() == ok ? 1 : 0;
}

int agwn_frac_batch_f32(float const* pos, int64_t n, float const* seg7, int64_t nseg, float const* x0, float* out)
{
    return guarded([&] { return frac_host(pos, n, seg7, nseg, x0, out); });
}
int agwn_frac_batch_f64(double const* pos, int64_t n, double const* seg7, int64_t nseg, double const* x0, double* out)
{
    return guarded([&] { return frac_host(pos, n, seg7, nseg, x0, out); });
}
int agwn_frac_batch_f32_dev(float const* pos, int64_t n, float const* seg7, int64_t nseg, float const* x0, float* out,
                            agwn_stream stream)
{
    return guarded([&] { return frac_dev(pos, n, seg7, nseg, x0, out, hipStream_t(stream)); });
}
int agwn_frac_batch_f64_dev(double const* pos, int64_t n, double const* seg7, int64_t nseg, double const* x0,
                            double* out, agwn_stream stream)
{
    return guarded([&] { return frac_dev(pos, n, seg7, nseg, x0, out, hipStream_t(stream)); });
}

int agwn_boundary_f32(float const* verts, int64_t nv, int32_t const* idx, int64_t nidx, float* out7, int64_t cap,
                      int64_t* count)
{
    return guarded([&] { return boundary_host(verts, nv, idx, nidx, out7, cap, count); });
}
int agwn_boundary_f64(double const* verts, int64_t nv, int32_t const* idx, int64_t nidx, double* out7, int64_t cap,
                      int64_t* count)
{
    return guarded([&] { return boundary_host(verts, nv, idx, nidx, out7, cap, count); });
}
} // extern "C"
