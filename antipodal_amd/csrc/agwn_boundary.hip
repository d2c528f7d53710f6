// Open boundary of a triangle mesh on the GPU: build_boundary_segments (math/common.hh:123-174) as a sort.
//
// Each directed half-edge u -> v of a triangle becomes the 64-bit key ((lo << 32 | hi) << 1) | (u < v), with
// lo = min(u, v), hi = max(u, v). An LSD radix sort (8-bit digits, 8 passes, LDS histograms and a stable in-tile
// ranking by wave ballots) groups the half-edges of each undirected edge; a segment-start pass sums the directions of
// each group (net = #(lo -> hi) - #(hi -> lo)); edges with net != 0 are compacted, in key order, into weighted
// segments oriented so that the weight |net| is positive. The multiset is the reference's; the order (by
// (lo, hi)) is the host implementation's, so the two agree element for element.

#include "../../include/agwn.h"
#include "agwn_internal.hpp"

#include <algorithm>

namespace agwn
{
namespace
{
constexpr int kB = 256;
constexpr int kItems = 16;
constexpr int kTileN = kB * kItems;

// Half-edge keys; flags an index outside [0, nv).
__global__ void __launch_bounds__(kB) k_edge_keys(int32_t const* __restrict__ idx, int64_t ntri, int64_t nv,
                                                  uint64_t* __restrict__ keys, int* __restrict__ bad)
{
    int64_t const h = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (h >= 3 * ntri)
        return;
    int64_t const t = h / 3;
    int const e = int(h % 3);
    int32_t const u = idx[3 * t + e], v = idx[3 * t + (e + 1) % 3];
    if (u < 0 || u >= nv || v < 0 || v >= nv)
    {
        atomicOr(bad, 1);
        keys[h] = 0;
        return;
    }
    uint32_t const lo = uint32_t(min(u, v)), hi = uint32_t(max(u, v));
    keys[h] = ((uint64_t(lo) << 32 | hi) << 1) | (u < v ? 1u : 0u);
}

__global__ void __launch_bounds__(kB) k_hist64(uint64_t const* __restrict__ keys, int64_t n, int shift,
                                               uint32_t* __restrict__ hist, int nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t const base = int64_t(blockIdx.x) * kTileN;
    for (int r = 0; r < kItems; ++r)
    {
        int64_t const i = base + int64_t(r) * kB + threadIdx.x;
        if (i < n)
            atomicAdd(&h[uint32_t(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m uint32 in place, one block of 1024 lanes (each lane scans a contiguous chunk).
__global__ void __launch_bounds__(1024) k_scan1024(uint32_t* __restrict__ a, int64_t m)
{
    __shared__ uint32_t part[1024];
    int64_t const chunk = (m + 1023) / 1024;
    int64_t const b = int64_t(threadIdx.x) * chunk;
    int64_t const e = b + chunk < m ? b + chunk : m;
    uint32_t s = 0;
    for (int64_t i = b; i < e; ++i)
        s += a[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 1; w < 1024; w <<= 1)
    {
        uint32_t const v = int(threadIdx.x) >= w ? part[threadIdx.x - w] : 0u;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (int64_t i = b; i < e; ++i)
    {
        uint32_t const v = a[i];
        a[i] = run;
        run += v;
    }
}

// Digit-pass offsets in two parallel levels: each digit's row of per-block counts (hist[d * nblocks + b]) is scanned by
// its own block in coalesced chunks of kB (wave shuffles, then the waves' sums), rowsum[d] = the row's total; then
// the 256 row totals are scanned by one block. Offset of (d, b) = hist[d * nblocks + b] + rowsum[d].
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* ws, uint32_t& total)
{
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1)
    {
        uint32_t const y = __shfl_up(x, o);
        if (lane >= o)
            x += y;
    }
    if (lane == 63)
        ws[wave] = x;
    __syncthreads();
    uint32_t base = 0;
    total = 0;
    for (int w = 0; w < kB / 64; ++w)
    {
        uint32_t const t = ws[w];
        if (w < wave)
            base += t;
        total += t;
    }
    __syncthreads(); // ws may be reused
    return base + x - v;
}
__global__ void __launch_bounds__(kB) k_scan_rows(uint32_t* __restrict__ hist, int nblocks,
                                                  uint32_t* __restrict__ rowsum)
{
    __shared__ uint32_t ws[kB / 64];
    uint32_t* row = hist + int64_t(blockIdx.x) * nblocks;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < nblocks; c0 += kB)
    {
        int const i = c0 + int(threadIdx.x);
        uint32_t const v = i < nblocks ? row[i] : 0u;
        uint32_t total = 0;
        uint32_t const e = block_scan_excl(v, ws, total);
        if (i < nblocks)
            row[i] = carry + e;
        carry += total;
    }
    if (threadIdx.x == 0)
        rowsum[blockIdx.x] = carry;
}
__global__ void __launch_bounds__(kB) k_scan_digits(uint32_t* __restrict__ rowsum)
{
    __shared__ uint32_t ws[kB / 64];
    uint32_t total = 0;
    uint32_t const e = block_scan_excl(rowsum[threadIdx.x], ws, total);
    rowsum[threadIdx.x] = e;
}

// Stable scatter of one digit pass. Each wave ranks its own contiguous chunk of kItems rows of 64 keys: lanes with
// equal digits are matched by ballots, and the wave's running per-digit counts live in LDS (the lowest lane of each
// match group advances them). The block then orders its waves per digit, places the tile's keys in LDS in digit
// order (stable: wave, row, lane is index order) and writes the tile out in that order, so that each digit's run of
// keys goes to consecutive addresses.
__global__ void __launch_bounds__(kB) k_scatter64(uint64_t const* __restrict__ kin, int64_t n, int shift,
                                                  uint32_t const* __restrict__ offs,
                                                  uint32_t const* __restrict__ rowbase, int nblocks,
                                                  uint64_t* __restrict__ kout)
{
    constexpr int kWaves = kB / 64;
    constexpr int kChunk = kItems * 64;
    static_assert(kB == 256, "one thread per digit");
    __shared__ uint64_t tile[kTileN];
    __shared__ uint32_t wc[kWaves][256]; // a wave's running count of each digit, then its offset within the digit
    __shared__ uint32_t lstart[256];     // the digit's first slot in the tile
    __shared__ uint32_t gdelta[256];     // the digit's first output index minus lstart
    __shared__ uint32_t ws[kWaves];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
        wc[w][threadIdx.x] = 0;
    __syncthreads();
    uint64_t const lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t const base = int64_t(blockIdx.x) * kTileN + int64_t(wave) * kChunk;
    uint64_t key[kItems];
    uint32_t rank[kItems]; // within the wave's keys of the same digit
#pragma unroll
    for (int r = 0; r < kItems; ++r)
    {
        int64_t const i = base + int64_t(r) * 64 + lane;
        bool const valid = i < n;
        key[r] = valid ? kin[i] : 0ull;
        uint32_t const d = uint32_t(key[r] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            bool const bit = (d >> b) & 1u;
            uint64_t const bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        // every lane reads before any leader writes: one LDS read and one write instruction of the same wave
        uint32_t const before = wc[wave][d];
        rank[r] = before + uint32_t(__popcll(m & lt));
        if (valid && (m & lt) == 0)
            wc[wave][d] = before + uint32_t(__popcll(m));
    }
    __syncthreads();
    // thread t = digit t: the waves' offsets within the digit, the digit's count and its start in the tile
    uint32_t cnt = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w)
    {
        uint32_t const c = wc[w][threadIdx.x];
        wc[w][threadIdx.x] = cnt;
        cnt += c;
    }
    uint32_t total = 0;
    uint32_t const start = block_scan_excl(cnt, ws, total);
    lstart[threadIdx.x] = start;
    gdelta[threadIdx.x] = offs[int64_t(threadIdx.x) * nblocks + blockIdx.x] + rowbase[threadIdx.x] - start;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kItems; ++r)
        if (base + int64_t(r) * 64 + lane < n)
        {
            uint32_t const d = uint32_t(key[r] >> shift) & 255u;
            tile[lstart[d] + wc[wave][d] + rank[r]] = key[r];
        }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += kB)
    {
        uint64_t const k = tile[j];
        uint32_t const pos = gdelta[uint32_t(k >> shift) & 255u] + j;
        if (pos < uint32_t(n)) // always true when the histogram matches; guards memory regardless
            kout[pos] = k;
    }
}

// flag[i] = 1 when sorted key i starts a group (undirected edge) whose net direction count is nonzero
__global__ void __launch_bounds__(kB) k_edge_net(uint64_t const* __restrict__ keys, int64_t n,
                                                 uint32_t* __restrict__ flag, int32_t* __restrict__ net)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n)
        return;
    uint64_t const e = keys[i] >> 1;
    bool const start = i == 0 || (keys[i - 1] >> 1) != e;
    int32_t s = 0;
    if (start)
        for (int64_t j = i; j < n && (keys[j] >> 1) == e; ++j)
            s += (keys[j] & 1u) ? 1 : -1;
    flag[i] = (start && s != 0) ? 1u : 0u;
    net[i] = s;
}

template <class T>
__global__ void __launch_bounds__(kB) k_emit_segments(uint64_t const* __restrict__ keys, int64_t n,
                                                      uint32_t const* __restrict__ flag_in,
                                                      uint32_t const* __restrict__ pos, int32_t const* __restrict__ net,
                                                      T const* __restrict__ verts, int64_t cap, T* __restrict__ out7)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n || !flag_in[i])
        return;
    int64_t const o = pos[i];
    if (o >= cap)
        return;
    uint64_t const e = keys[i] >> 1;
    int64_t const lo = int64_t(e >> 32), hi = int64_t(e & 0xffffffffu);
    int32_t const s = net[i];
    int64_t const a = s > 0 ? lo : hi, b = s > 0 ? hi : lo;
    T* q = out7 + 7 * o;
    for (int k = 0; k < 3; ++k)
    {
        q[k] = verts[3 * a + k];
        q[3 + k] = verts[3 * b + k];
    }
    q[6] = T(s > 0 ? s : -s);
}

template <class T>
int boundary_dev(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T* out7, int64_t cap, int64_t* count,
                 hipStream_t s)
{
    AGWN_CHECK(nv >= 0 && nidx >= 0 && cap >= 0, "negative count");
    AGWN_CHECK(nidx % 3 == 0, "index count %lld is not a multiple of 3", (long long)nidx);
    AGWN_CHECK(count != nullptr, "count is null");
    AGWN_CHECK(nidx == 0 || (verts && idx), "null mesh arrays");
    AGWN_CHECK(cap == 0 || out7, "out7 is null with cap=%lld", (long long)cap);
    AGWN_CHECK(nidx < (int64_t(1) << 31), "too many indices (%lld)", (long long)nidx);
    *count = 0;
    if (nidx == 0)
        return ok;
    AGWN_TRY(ensure_device());
    int64_t const n = nidx; // half-edges
    dbuf k0, k1, flag, net, bad;
    AGWN_TRY(k0.alloc(size_t(n) * sizeof(uint64_t), s));
    AGWN_TRY(k1.alloc(size_t(n) * sizeof(uint64_t), s));
    AGWN_TRY(flag.alloc(size_t(n + 1) * sizeof(uint32_t), s));
    AGWN_TRY(net.alloc(size_t(n) * sizeof(int32_t), s));
    AGWN_TRY(bad.alloc(sizeof(int), s));
    AGWN_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    unsigned const g = unsigned((n + kB - 1) / kB);
    k_edge_keys<<<dim3(g), dim3(kB), 0, s>>>(idx, nidx / 3, nv, k0.as<uint64_t>(), bad.as<int>());
    AGWN_HIP(hipGetLastError());
    int hbad = 0;
    AGWN_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    AGWN_HIP(hipStreamSynchronize(s));
    AGWN_CHECK(hbad == 0, "a triangle index is out of range [0, %lld)", (long long)nv);
    // the key's significant bits: 2 * 31 + 1 = 63 (indices < 2^31)
    uint64_t* ka = nullptr;
    AGWN_TRY(radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), n, 64, s, &ka));
    // ka holds the sorted keys; flags -> exclusive positions (flag[n] = total after the scan)
    uint32_t* fl = flag.as<uint32_t>();
    k_edge_net<<<dim3(g), dim3(kB), 0, s>>>(ka, n, fl, net.as<int32_t>());
    AGWN_HIP(hipMemsetAsync(fl + n, 0, sizeof(uint32_t), s));
    dbuf pos;
    AGWN_TRY(pos.alloc(size_t(n + 1) * sizeof(uint32_t), s));
    AGWN_HIP(hipMemcpyAsync(pos.p, fl, size_t(n + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    k_scan1024<<<dim3(1), dim3(1024), 0, s>>>(pos.as<uint32_t>(), n + 1);
    AGWN_HIP(hipGetLastError());
    uint32_t total = 0;
    AGWN_HIP(hipMemcpyAsync(&total, pos.as<uint32_t>() + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (cap > 0)
    {
        k_emit_segments<T><<<dim3(g), dim3(kB), 0, s>>>(ka, n, fl, pos.as<uint32_t>(), net.as<int32_t>(), verts, cap,
                                                        out7);
        AGWN_HIP(hipGetLastError());
    }
    AGWN_HIP(hipStreamSynchronize(s));
    *count = int64_t(total);
    return ok;
}

} // namespace

// Scratch of radix_sort_u64_ws for n keys, in uint32 words: the per-tile digit histograms and the 256 digit starts.
int64_t radix_sort_words(int64_t n) { return 256 * ((n + kTileN - 1) / kTileN) + 256; }

// LSD radix sort of n uint64 keys over their low `bits` bits (8-bit digits) with the caller's scratch
// (radix_sort_words(n) uint32); only kernel launches, so it can be captured in a graph; *sorted points at a or b.
int radix_sort_u64_ws(uint64_t* a, uint64_t* b, int64_t n, int bits, uint32_t* ws, hipStream_t s, uint64_t** sorted)
{
    *sorted = a;
    if (n <= 1)
        return ok;
    AGWN_CHECK(n < (int64_t(1) << 32), "radix sort of %lld keys", (long long)n);
    int const nblocks = int((n + kTileN - 1) / kTileN);
    uint32_t* const hist = ws;
    uint32_t* const rows = ws + size_t(256) * nblocks;
    for (int shift = 0; shift < bits; shift += 8)
    {
        k_hist64<<<dim3(unsigned(nblocks)), dim3(kB), 0, s>>>(a, n, shift, hist, nblocks);
        k_scan_rows<<<dim3(256), dim3(kB), 0, s>>>(hist, nblocks, rows);
        k_scan_digits<<<dim3(1), dim3(kB), 0, s>>>(rows);
        k_scatter64<<<dim3(unsigned(nblocks)), dim3(kB), 0, s>>>(a, n, shift, hist, rows, nblocks, b);
        AGWN_HIP(hipGetLastError());
        std::swap(a, b);
    }
    *sorted = a;
    return ok;
}

// The same with scratch of its own; *sorted points at a or b.
int radix_sort_u64(uint64_t* a, uint64_t* b, int64_t n, int bits, hipStream_t s, uint64_t** sorted)
{
    *sorted = a;
    if (n <= 1)
        return ok;
    dbuf ws;
    AGWN_TRY(ws.alloc(size_t(radix_sort_words(n)) * sizeof(uint32_t), s));
    return radix_sort_u64_ws(a, b, n, bits, ws.as<uint32_t>(), s, sorted);
}

namespace
{
template <class F>
int guarded_b(F&& f)
{
    clear_error();
    try
    {
        return f();
    }
    catch (...)
    {
        set_error("unexpected C++ exception");
        return err_hip;
    }
}
} // namespace
} // namespace agwn

extern "C" {
int agwn_boundary_f32_dev(float const* verts, int64_t nv, int32_t const* idx, int64_t nidx, float* out7, int64_t cap,
                          int64_t* count, agwn_stream s)
{
    return agwn::guarded_b([&] { return agwn::boundary_dev(verts, nv, idx, nidx, out7, cap, count, hipStream_t(s)); });
}
int agwn_boundary_f64_dev(double const* verts, int64_t nv, int32_t const* idx, int64_t nidx, double* out7,
                          int64_t cap, int64_t* count, agwn_stream s)
{
    return agwn::guarded_b([&] { return agwn::boundary_dev(verts, nv, idx, nidx, out7, cap, count, hipStream_t(s)); });
}
} // extern "C"
