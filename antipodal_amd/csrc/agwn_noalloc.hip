// Allocation-free FP64 fractional term: agwn_frac_batch_f64_dev without any device allocation, for callers that
// cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The FP64 fractional kernels need one workspace: the boundary rotated into the pole frame of x0, one 64-byte seg_f64
// per segment. Here it is the file-scope device array `scratch` (kNoAllocMaxSeg segments = 32 MiB), which every GPU
// the library runs on holds from module load until the process ends. Its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) on the entry point's first call there and kept per device: in host
// code the array's name is only a host-side stand-in, so only the looked-up address is ever passed to a kernel.
//
// Every call on one GPU shares that GPU's array, and the array keeps the last call's contents. Nothing in it must
// start at zero: the first kernel writes entries [0, nseg) and the second reads only those. Calls that overlap on one
// GPU (several host threads, or several streams without ordering between them) race on the array and must be
// serialised by the caller.

#include "agwn_internal.hpp"
#include "../../include/agwn.h"

#include <mutex>

namespace agwn
{
namespace
{
constexpr int64_t kNoAllocMaxSeg = int64_t(1) << 19; // 524,288 segments x 64 B = 32 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) seg_f64 scratch[kNoAllocMaxSeg];

std::mutex g_scratch_mu;
seg_f64* g_scratch[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
int scratch_address(seg_f64** out)
{
    int dev = -1;
    AGWN_HIP(hipGetDevice(&dev));
    AGWN_CHECK(dev >= 0 && dev < kMaxDevices, "device ordinal %d beyond the %d this library tracks", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev] == nullptr)
    {
        void* p = nullptr;
        AGWN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        AGWN_CHECK(p != nullptr, "hipGetSymbolAddress returned a null address for the scratch array");
        g_scratch[dev] = static_cast<seg_f64*>(p);
    }
    *out = g_scratch[dev];
    return ok;
}

int frac_f64_noalloc(double const* pos, int64_t n, double const* seg7, int64_t nseg, double const* x0, double* out,
                     hipStream_t s)
{
    AGWN_CHECK(n >= 0 && nseg >= 0, "negative count (n=%lld, nseg=%lld)", (long long)n, (long long)nseg);
    AGWN_CHECK(nseg <= kNoAllocMaxSeg,
               "boundary of %lld segments exceeds the allocation-free entry point's maximum of %lld segments (use "
               "agwn_frac_batch_f64_dev)",
               (long long)nseg, (long long)kNoAllocMaxSeg);
    AGWN_CHECK(x0 != nullptr, "x0 is null");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    if (n == 0)
        return ok;
    AGWN_TRY(ensure_device());
    seg_f64* ws = nullptr;
    AGWN_TRY(scratch_address(&ws));
    double const x[3] = {x0[0], x0[1], x0[2]};
    return frac_f64_ws(src_array<double>{pos}, n, seg7, nseg, x, nullptr, out, ws, s);
}
} // namespace
} // namespace agwn

extern "C" int agwn_frac_batch_f64_dev_noalloc(double const* pos, int64_t n, double const* seg7, int64_t nseg,
                                               double const* x0, double* out, agwn_stream stream)
{
    agwn::clear_error();
    try
    {
        return agwn::frac_f64_noalloc(pos, n, seg7, nseg, x0, out, hipStream_t(stream));
    }
    catch (...)
    {
        agwn::set_error("unexpected C++ exception");
        return agwn::err_invalid;
    }
}
