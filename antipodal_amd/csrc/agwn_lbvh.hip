// LBVH construction on gfx950: Morton codes, hand-written LDS radix sort, Karras hierarchy, atomic bottom-up refit.

#include "agwn_lbvh.hpp"

#include <algorithm>
#include <vector>

namespace agwn
{
namespace
{
constexpr int kB = 256;

// order-preserving float <-> uint encoding for atomicMin/atomicMax
__device__ __forceinline__ uint32_t fenc(float f)
{
    uint32_t const u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fdec(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

__global__ void __launch_bounds__(kB) k_bounds(float const* __restrict__ c, int64_t n, uint32_t* __restrict__ b6)
{
    __shared__ float smin[3][kB], smax[3][kB];
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
        for (int k = 0; k < 3; ++k)
        {
            float const v = c[3 * i + k];
            mn[k] = fminf(mn[k], v);
            mx[k] = fmaxf(mx[k], v);
        }
    for (int k = 0; k < 3; ++k)
    {
        smin[k][threadIdx.x] = mn[k];
        smax[k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w)
            for (int k = 0; k < 3; ++k)
            {
                smin[k][threadIdx.x] = fminf(smin[k][threadIdx.x], smin[k][threadIdx.x + w]);
                smax[k][threadIdx.x] = fmaxf(smax[k][threadIdx.x], smax[k][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k)
        {
            atomicMin(&b6[k], fenc(smin[k][0]));
            atomicMax(&b6[3 + k], fenc(smax[k][0]));
        }
}

__device__ __forceinline__ uint32_t expand10(uint32_t v)
{
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void __launch_bounds__(kB) k_morton(float const* __restrict__ c, int64_t n, uint32_t const* __restrict__ b6,
                                               uint32_t* __restrict__ keys, int32_t* __restrict__ vals)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n)
        return;
    uint32_t code = 0;
    for (int k = 0; k < 3; ++k)
    {
        float const lo = fdec(b6[k]), hi = fdec(b6[3 + k]);
        float const ext = hi - lo;
        float t = ext > 0.0f ? (c[3 * i + k] - lo) / ext : 0.5f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        uint32_t const q = uint32_t(fminf(t * 1024.0f, 1023.0f));
        code |= expand10(q) << (2 - k);
    }
    keys[i] = code;
    vals[i] = int32_t(i);
}

// ---------------------------------------------------------------------------------------------------------------------
// LSD radix sort of (uint32 key, int32 value), 8-bit digits. Tile = 256 lanes x 16 items.
// ---------------------------------------------------------------------------------------------------------------------
constexpr int kRsItems = 16;
constexpr int kRsTile = kB * kRsItems;

__global__ void __launch_bounds__(kB) k_rs_hist(uint32_t const* __restrict__ keys, int64_t n, int shift,
                                                uint32_t* __restrict__ hist, int nblocks)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int64_t const base = int64_t(blockIdx.x) * kRsTile;
    for (int r = 0; r < kRsItems; ++r)
    {
        int64_t const i = base + int64_t(r) * kB + threadIdx.x;
        if (i < n)
            atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[int64_t(threadIdx.x) * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of m uint32 in place, one block of 1024 lanes (each lane scans a contiguous chunk).
__global__ void __launch_bounds__(1024) k_scan_excl(uint32_t* __restrict__ a, int64_t m)
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
    // Hillis-Steele inclusive scan over the 1024 partial sums
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

__global__ void __launch_bounds__(kB) k_rs_scatter(uint32_t const* __restrict__ kin, int32_t const* __restrict__ vin,
                                                   int64_t n, int shift, uint32_t const* __restrict__ offs,
                                                   int nblocks, uint32_t* __restrict__ kout,
                                                   int32_t* __restrict__ vout)
{
    constexpr int kWaves = kB / 64;
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[kWaves][256];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    run[threadIdx.x] = offs[int64_t(threadIdx.x) * nblocks + blockIdx.x];
    uint64_t const lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int64_t const base = int64_t(blockIdx.x) * kRsTile;
    for (int r = 0; r < kRsItems; ++r)
    {
        int64_t const i = base + int64_t(r) * kB + threadIdx.x;
        bool const valid = i < n;
        uint32_t const key = valid ? kin[i] : 0u;
        int32_t const val = valid ? vin[i] : 0;
        uint32_t const d = (key >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b)
        {
            bool const bit = (d >> b) & 1u;
            uint64_t const bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        for (int w = 0; w < kWaves; ++w)
            wcnt[w][threadIdx.x] = 0;
        __syncthreads();
        bool const leader = valid && (m & lt) == 0;
        if (leader)
            wcnt[wave][d] = uint32_t(__popcll(m));
        __syncthreads();
        if (valid)
        {
            uint32_t pos = run[d] + uint32_t(__popcll(m & lt));
            for (int w = 0; w < wave; ++w)
                pos += wcnt[w][d];
            if (pos < uint32_t(n)) // always true when the histogram matches; guards memory regardless
            {
                kout[pos] = key;
                vout[pos] = val;
            }
        }
        __syncthreads();
        uint32_t tot = 0;
        for (int w = 0; w < kWaves; ++w)
            tot += wcnt[w][threadIdx.x];
        run[threadIdx.x] += tot;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// Karras (2012) hierarchy. delta(i, j) = common prefix length of the keys, extended by the positions on ties.
// ---------------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ int delta(uint32_t const* __restrict__ k, int64_t n, int64_t i, int64_t j)
{
    if (j < 0 || j >= n)
        return -1;
    uint32_t const a = k[i], b = k[j];
    if (a != b)
        return __clz(a ^ b);
    return 32 + __clz(uint32_t(i ^ j)); // positions are < 2^31
}

__global__ void __launch_bounds__(kB) k_karras(uint32_t const* __restrict__ k, int64_t n, int32_t* __restrict__ child,
                                               int32_t* __restrict__ parent, int8_t* __restrict__ big,
                                               uint32_t* __restrict__ csz)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n - 1)
        return;
    int const d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
    int const dmin = delta(k, n, i, i - d);
    int64_t lmax = 2;
    while (delta(k, n, i, i + lmax * d) > dmin)
        lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(k, n, i, i + (l + t) * d) > dmin)
            l += t;
    int64_t const j = i + l * d;
    int const dnode = delta(k, n, i, j);
    int64_t s = 0;
    for (int64_t t = (l + 1) / 2;; t = (t + 1) / 2)
    {
        if (delta(k, n, i, i + (s + t) * d) > dnode)
            s += t;
        if (t == 1)
            break;
    }
    int64_t const g = i + s * d + (d < 0 ? -1 : 0);
    int64_t const lo = i < j ? i : j, hi = i < j ? j : i;
    int32_t const left = (lo == g) ? ~int32_t(g) : int32_t(g);
    int32_t const right = (hi == g + 1) ? ~int32_t(g + 1) : int32_t(g + 1);
    child[2 * i] = left;
    child[2 * i + 1] = right;
    big[i] = (hi - g) > (g - lo + 1) ? 1 : 0; // leaves right: [g+1, hi], left: [lo, g]
    csz[i] = uint32_t(g - lo + 1 < 255 ? g - lo + 1 : 255) | (uint32_t(hi - g < 255 ? hi - g : 255) << 8);
    parent[left >= 0 ? left : (n - 1) + ~left] = int32_t(i);
    parent[right >= 0 ? right : (n - 1) + ~right] = int32_t(i);
}

__global__ void __launch_bounds__(kB) k_depth(int32_t const* __restrict__ parent, int64_t n, int* __restrict__ maxd)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= n)
        return;
    int d = 0;
    int32_t p = parent[(n - 1) + k];
    while (p >= 0 && d < 4096)
    {
        ++d;
        p = parent[p];
    }
    atomicMax(maxd, d);
}

template <class B>
__device__ __forceinline__ void ld_box(B const* lo, B const* hi, int64_t i, B (&l)[3], B (&h)[3])
{
    for (int k = 0; k < 3; ++k)
    {
        l[k] = __hip_atomic_load(const_cast<B*>(lo) + 3 * i + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        h[k] = __hip_atomic_load(const_cast<B*>(hi) + 3 * i + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class B>
__global__ void __launch_bounds__(kB)
    k_refit(int32_t const* __restrict__ child, int32_t const* __restrict__ parent, int8_t const* __restrict__ big,
            uint32_t const* __restrict__ csz, int64_t n, B const* leaf_lo, B const* leaf_hi, B* node_lo, B* node_hi, bvh_node<B>* __restrict__ nodes,
            int32_t* flags)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= n)
        return;
    int32_t p = parent[(n - 1) + k];
    while (p >= 0)
    {
        __threadfence();
        if (atomicAdd(&flags[p], 1) == 0)
            return; // the sibling subtree is not finished; its last thread continues
        __threadfence();
        int32_t const c[2] = {child[2 * p], child[2 * p + 1]};
        B l[2][3], h[2][3];
        for (int s = 0; s < 2; ++s)
        {
            if (c[s] < 0)
                ld_box(leaf_lo, leaf_hi, ~c[s], l[s], h[s]);
            else
                ld_box<B>(node_lo, node_hi, c[s], l[s], h[s]);
        }
        bvh_node<B> nd;
        for (int a = 0; a < 3; ++a)
        {
            nd.lo0[a] = l[0][a];
            nd.hi0[a] = h[0][a];
            nd.lo1[a] = l[1][a];
            nd.hi1[a] = h[1][a];
            B const ul = l[0][a] < l[1][a] ? l[0][a] : l[1][a];
            B const uh = h[0][a] > h[1][a] ? h[0][a] : h[1][a];
            __hip_atomic_store(node_lo + 3 * p + a, ul, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(node_hi + 3 * p + a, uh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        nd.c0 = c[0];
        nd.c1 = c[1];
        nd.big = big[p];
        nd.sz = csz[p];
        nodes[p] = nd;
        p = parent[p];
    }
}
} // namespace

int lbvh_build_topology(float const* centroids, int64_t n, lbvh_topology& topo, hipStream_t s)
{
    lbvh_release(topo);
    topo.n = n;
    if (n <= 0)
        return ok;
    AGWN_CHECK(n < (int64_t(1) << 27), "too many triangles for the LBVH (%lld; the traversal stack holds 2^27)",
               (long long)n);
    AGWN_HIP(hipMalloc(&topo.order, size_t(n) * sizeof(int32_t)));
    AGWN_HIP(hipMalloc(&topo.parent, size_t(2 * n) * sizeof(int32_t)));
    AGWN_HIP(hipMalloc(&topo.child, size_t(std::max<int64_t>(2 * (n - 1), 1)) * sizeof(int32_t)));
    AGWN_HIP(hipMalloc(&topo.big, size_t(std::max<int64_t>(n - 1, 1)) * sizeof(int8_t)));
    AGWN_HIP(hipMalloc(&topo.csz, size_t(std::max<int64_t>(n - 1, 1)) * sizeof(uint32_t)));
    if (n == 1)
    {
        int32_t const zero = 0, none = -1;
        AGWN_HIP(hipMemcpyAsync(topo.order, &zero, sizeof(int32_t), hipMemcpyHostToDevice, s));
        AGWN_HIP(hipMemcpyAsync(topo.parent, &none, sizeof(int32_t), hipMemcpyHostToDevice, s));
        AGWN_HIP(hipStreamSynchronize(s));
        topo.max_depth = 0;
        return ok;
    }
    dbuf b6, k0, k1, v1, hist, md;
    int const nblocks = int((n + kRsTile - 1) / kRsTile);
    AGWN_TRY(b6.alloc(6 * sizeof(uint32_t), s));
    AGWN_TRY(k0.alloc(size_t(n) * sizeof(uint32_t), s));
    AGWN_TRY(k1.alloc(size_t(n) * sizeof(uint32_t), s));
    AGWN_TRY(v1.alloc(size_t(n) * sizeof(int32_t), s));
    AGWN_TRY(hist.alloc(size_t(256) * nblocks * sizeof(uint32_t), s));
    AGWN_TRY(md.alloc(sizeof(int), s));
    uint32_t const init[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    AGWN_HIP(hipMemcpyAsync(b6.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    unsigned const g = unsigned((n + kB - 1) / kB);
    k_bounds<<<dim3(std::min(g, 1024u)), dim3(kB), 0, s>>>(centroids, n, b6.as<uint32_t>());
    AGWN_HIP(hipGetLastError());
    // keys/values ping-pong: (k0, order) -> (k1, v1) -> (k0, order) ... 4 passes end in (k0, order)
    k_morton<<<dim3(g), dim3(kB), 0, s>>>(centroids, n, b6.as<uint32_t>(), k0.as<uint32_t>(), topo.order);
    AGWN_HIP(hipGetLastError());
    uint32_t* ka = k0.as<uint32_t>();
    uint32_t* kb = k1.as<uint32_t>();
    int32_t* va = topo.order;
    int32_t* vb = v1.as<int32_t>();
    for (int shift = 0; shift < 32; shift += 8)
    {
        k_rs_hist<<<dim3(unsigned(nblocks)), dim3(kB), 0, s>>>(ka, n, shift, hist.as<uint32_t>(), nblocks);
        k_scan_excl<<<dim3(1), dim3(1024), 0, s>>>(hist.as<uint32_t>(), int64_t(256) * nblocks);
        k_rs_scatter<<<dim3(unsigned(nblocks)), dim3(kB), 0, s>>>(ka, va, n, shift, hist.as<uint32_t>(), nblocks, kb,
                                                                 vb);
        AGWN_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // after 4 passes the sorted data is back in (k0, order)
    AGWN_HIP(hipMemsetAsync(topo.parent, 0xff, size_t(2 * n) * sizeof(int32_t), s));
    k_karras<<<dim3(unsigned((n - 1 + kB - 1) / kB)), dim3(kB), 0, s>>>(ka, n, topo.child, topo.parent, topo.big,
                                                                                topo.csz);
    AGWN_HIP(hipGetLastError());
    AGWN_HIP(hipMemsetAsync(md.p, 0, sizeof(int), s));
    k_depth<<<dim3(g), dim3(kB), 0, s>>>(topo.parent, n, md.as<int>());
    AGWN_HIP(hipGetLastError());
    int depth = 0;
    AGWN_HIP(hipMemcpyAsync(&depth, md.p, sizeof(int), hipMemcpyDeviceToHost, s));
    AGWN_HIP(hipStreamSynchronize(s));
    topo.max_depth = depth;
    return ok;
}

void lbvh_release(lbvh_topology& topo)
{
    if (topo.order)
        (void)hipFree(topo.order);
    if (topo.child)
        (void)hipFree(topo.child);
    if (topo.parent)
        (void)hipFree(topo.parent);
    if (topo.big)
        (void)hipFree(topo.big);
    if (topo.csz)
        (void)hipFree(topo.csz);
    topo = lbvh_topology{};
}

template <class B>
int lbvh_refit(lbvh_topology const& topo, B const* leaf_lo, B const* leaf_hi, bvh_node<B>* nodes, B* node_lo,
               B* node_hi, int32_t* flags, hipStream_t s)
{
    int64_t const n = topo.n;
    if (n <= 1)
        return ok;
    AGWN_HIP(hipMemsetAsync(flags, 0, size_t(n - 1) * sizeof(int32_t), s));
    k_refit<B><<<dim3(unsigned((n + kB - 1) / kB)), dim3(kB), 0, s>>>(topo.child, topo.parent, topo.big, topo.csz, n,
                                                                     leaf_lo, leaf_hi, node_lo, node_hi, nodes, flags);
    AGWN_HIP(hipGetLastError());
    return ok;
}

template int lbvh_refit<float>(lbvh_topology const&, float const*, float const*, bvh_node<float>*, float*, float*,
                               int32_t*, hipStream_t);
template int lbvh_refit<double>(lbvh_topology const&, double const*, double const*, bvh_node<double>*, double*,
                                double*, int32_t*, hipStream_t);
} // namespace agwn
