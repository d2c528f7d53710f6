// Prepared mesh (agwn_mesh) and the plain integer term: signed ray-crossing counts along an arbitrary direction,
// bit-identical to NaiveIntersector<T>::signed_intersection_count (intersector/intersector.hh:61-100).
//
// Exactness argument. The reference tests every triangle with Moeller-Trumbore in T. Here the per-direction part of
// that test (e1 = b - a, e2 = c - a, h = d x e2, det = e1 . h, inv = 1/det, sign(d . (e1 x e2))) is computed once per
// call per triangle with the same operations (no contraction, IEEE division), and the per-query part (s = p - a,
// u, q = s x e1, v, t) per candidate with the same operations -- so every tested triangle gives the reference's
// verdict bit for bit. The LBVH only decides which triangles are tested. It must never cull a triangle the FP test
// would accept: each leaf box is the triangle's AABB grown by pad = 32 eps_T (max|p| + max|v|) cond, with
// cond = |e1||e2||d| / |det|, which bounds how far from the triangle the rounded test can accept a ray (the rounding of
// s, of the three dot/cross products and of the multiplications by inv, each O(eps |s| |e| / |det|) in barycentric
// units), rounded outward to float; the float slab test is made conservative by scaling t_far by 1 + 2 gamma_3 (Ize,
// "Robust BVH Ray Traversal", JCGT 2013). Triangles with det == 0 are never accepted by the reference and are dropped;
// triangles with cond above 2^13 (f32) / 2^40 (f64) -- nearly parallel to d -- would need boxes so large that they
// are tested exhaustively by every query instead (the "brute" list).

#include "../../include/agwn.h"
#include "agwn_lbvh.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <mutex>
#include <type_traits>

namespace agwn
{
namespace
{
constexpr int kB = 256;
constexpr int kTB = 128; // traversal block: per-lane stacks (count_one) take kMaxStack x 4 B x 128 lanes = 14 KiB of LDS

template <class T>
struct alignas(sizeof(T) == 4 ? 64 : 128) leaf_rec
{
    T ax, ay, az, e1x, e1y, e1z, e2x, e2y, e2z, hx, hy, hz, inv;
    int32_t sign;
    int32_t live; // 1: tested through the BVH; 0: never hit (det == 0) or in the brute list
};

template <class T>
__global__ void __launch_bounds__(kB) k_centroids(T const* __restrict__ v, int32_t const* __restrict__ idx, int64_t ntri,
                                                  float* __restrict__ c)
{
    int64_t const t = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (t >= ntri)
        return;
    for (int k = 0; k < 3; ++k)
    {
        double const s = double(v[3 * int64_t(idx[3 * t]) + k]) + double(v[3 * int64_t(idx[3 * t + 1]) + k]) +
                         double(v[3 * int64_t(idx[3 * t + 2]) + k]);
        c[3 * t + k] = float(s / 3.0);
    }
}

// sorted triangle vertices: tri[9*k..] = a, b, c of the triangle at sorted position k
template <class T>
__global__ void __launch_bounds__(kB) k_gather_tris(T const* __restrict__ v, int32_t const* __restrict__ idx,
                                                    int32_t const* __restrict__ order, int64_t ntri, T* __restrict__ tri)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= ntri)
        return;
    int64_t const t = order[k];
    for (int j = 0; j < 3; ++j)
        for (int c = 0; c < 3; ++c)
            tri[9 * k + 3 * j + c] = v[3 * int64_t(idx[3 * t + j]) + c];
}

// max over queries of |p|_2 (as double bits via atomicMax on the non-negative encoding)
template <class T>
__global__ void __launch_bounds__(kB) k_max_norm(T const* __restrict__ p, int64_t n, unsigned long long* __restrict__ out)
{
    __shared__ double red[kB];
    double m = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < n; i += int64_t(gridDim.x) * kB)
    {
        double const x = double(p[3 * i]), y = double(p[3 * i + 1]), z = double(p[3 * i + 2]);
        m = fmax(m, sqrt(x * x + y * y + z * z));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kB / 2; w > 0; w >>= 1)
    {
        if (int(threadIdx.x) < w)
            red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

__device__ __forceinline__ float round_down(double x)
{
    float f = float(x);
    if (double(f) > x)
        f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float round_up(double x)
{
    float f = float(x);
    if (double(f) < x)
        f = nextafterf(f, INFINITY);
    return f;
}

// Per-call, per-direction leaf preparation.
template <class T>
__global__ void __launch_bounds__(kB)
    k_prep_leaves(T const* __restrict__ tri, int64_t ntri, T dx, T dy, T dz, double dsum, double eps, double float_pad,
                  double cond_max, int force_brute, leaf_rec<T>* __restrict__ rec, float* __restrict__ lo,
                  float* __restrict__ hi, int32_t* __restrict__ brute, int32_t* __restrict__ nbrute)
{
#pragma clang fp contract(off)
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= ntri)
        return;
    T const* g = tri + 9 * k;
    T const ax = g[0], ay = g[1], az = g[2];
    T const bx = g[3], by = g[4], bz = g[5];
    T const cx = g[6], cy = g[7], cz = g[8];
    // intersector.hh:73-80, same operations in the same order
    T const e1x = bx - ax, e1y = by - ay, e1z = bz - az;
    T const e2x = cx - ax, e2y = cy - ay, e2z = cz - az;
    T const hx = dy * e2z - dz * e2y;
    T const hy = dz * e2x - dx * e2z;
    T const hz = dx * e2y - dy * e2x;
    T const det = e1x * hx + e1y * hy + e1z * hz;
    // intersector.hh:95-97: sign of (e1 x e2) . dir, ties to -1
    T const nx = e1y * e2z - e1z * e2y;
    T const ny = e1z * e2x - e1x * e2z;
    T const nz = e1x * e2y - e1y * e2x;
    T const nd = nx * dx + ny * dy + nz * dz;
    leaf_rec<T> r;
    r.ax = ax;
    r.ay = ay;
    r.az = az;
    r.e1x = e1x;
    r.e1y = e1y;
    r.e1z = e1z;
    r.e2x = e2x;
    r.e2y = e2y;
    r.e2z = e2z;
    r.hx = hx;
    r.hy = hy;
    r.hz = hz;
    r.inv = det != T(0) ? T(1) / det : T(0);
    r.sign = nd > T(0) ? 1 : -1;
    r.live = 0;
    float l[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, h[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX}; // empty
    if (det != T(0))
    {
        double const n1 = sqrt(double(e1x) * e1x + double(e1y) * e1y + double(e1z) * e1z);
        double const n2 = sqrt(double(e2x) * e2x + double(e2y) * e2y + double(e2z) * e2z);
        double const nd_ = sqrt(double(dx) * dx + double(dy) * dy + double(dz) * dz);
        double const cond = n1 * n2 * nd_ / fabs(double(det));
        if (cond > cond_max || force_brute)
        {
            brute[atomicAdd(nbrute, 1)] = int32_t(k);
        }
        else
        {
            r.live = 1;
            double const pad = 32.0 * eps * dsum * cond + float_pad;
            double const vx[3] = {double(ax), double(bx), double(cx)};
            double const vy[3] = {double(ay), double(by), double(cy)};
            double const vz[3] = {double(az), double(bz), double(cz)};
            double const mn[3] = {fmin(fmin(vx[0], vx[1]), vx[2]), fmin(fmin(vy[0], vy[1]), vy[2]),
                                  fmin(fmin(vz[0], vz[1]), vz[2])};
            double const mx[3] = {fmax(fmax(vx[0], vx[1]), vx[2]), fmax(fmax(vy[0], vy[1]), vy[2]),
                                  fmax(fmax(vz[0], vz[1]), vz[2])};
            for (int a = 0; a < 3; ++a)
            {
                l[a] = round_down(mn[a] - pad);
                h[a] = round_up(mx[a] + pad);
            }
        }
    }
    rec[k] = r;
    for (int a = 0; a < 3; ++a)
    {
        lo[3 * k + a] = l[a];
        hi[3 * k + a] = h[a];
    }
}

// Moeller-Trumbore, query part (intersector.hh:81-97), reference operation order, no contraction.
template <class T>
__device__ __forceinline__ int mt_test(leaf_rec<T> const& r, T px, T py, T pz, T dx, T dy, T dz)
{
#pragma clang fp contract(off)
    T const sx = px - r.ax, sy = py - r.ay, sz = pz - r.az;
    T const u = r.inv * (sx * r.hx + sy * r.hy + sz * r.hz);
    if (u < T(0) || u > T(1))
        return 0;
    T const qx = sy * r.e1z - sz * r.e1y;
    T const qy = sz * r.e1x - sx * r.e1z;
    T const qz = sx * r.e1y - sy * r.e1x;
    T const v = r.inv * (dx * qx + dy * qy + dz * qz);
    if (v < T(0) || u + v > T(1))
        return 0;
    T const t = r.inv * (r.e2x * qx + r.e2y * qy + r.e2z * qz);
    if (t < T(0))
        return 0;
    return r.sign;
}

// conservative float slab test of the ray (o, 1/d) against [lo, hi] for t >= 0
__device__ __forceinline__ bool slab(float const* lo, float const* hi, float ox, float oy, float oz, float ix, float iy,
                                     float iz)
{
    if (!(lo[0] <= hi[0]))
        return false; // empty box
    float const tx0 = (lo[0] - ox) * ix, tx1 = (hi[0] - ox) * ix;
    float const ty0 = (lo[1] - oy) * iy, ty1 = (hi[1] - oy) * iy;
    float const tz0 = (lo[2] - oz) * iz, tz1 = (hi[2] - oz) * iz;
    float const tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
    float const tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    return tn <= tf * 1.0000004f; // 1 + 2 gamma_3, gamma_3 = 3u/(1-3u), u = 2^-24
}

__device__ __forceinline__ float inv_dir(double d)
{
    float const f = float(d);
    return f != 0.0f ? 1.0f / f : (signbit(f) ? -1e30f : 1e30f);
}

// Signed crossing count of one ray (p, d) against a prepared mesh: LBVH short-stack traversal (stack column `lane` of
// an LDS array [kMaxStack][kTB]) with the Moeller-Trumbore replica at the leaves, plus the exhaustively tested
// triangles. The batched kernel's traversal: its queries are scattered, where a wave's rays share few nodes. With
// Buckets, internal children of at most L leaves are taken as leaf buckets (bucket_span): their leaves are tested
// without their own boxes. A leaf box only ever rejects triangles the test would not count (it holds every point where
// the replica can report a crossing), so the counts do not depend on L.
template <bool Buckets, class T>
__device__ __forceinline__ int count_one(T px, T py, T pz, T dx, T dy, T dz, bvh_node<float> const* __restrict__ nodes,
                                         leaf_rec<T> const* __restrict__ rec, int32_t const* __restrict__ brute,
                                         int32_t nb, int use_bvh, int L, int32_t (*stack)[kTB], int lane)
{
    int sum = 0;
    auto leaves = [&](int32_t first, int32_t count) __attribute__((always_inline)) {
        for (int32_t k = first; k < first + count; ++k)
        {
            leaf_rec<T> const r = rec[k];
            if (r.live)
                sum += mt_test(r, px, py, pz, dx, dy, dz);
        }
    };
    if (use_bvh)
    {
        float const ox = float(px), oy = float(py), oz = float(pz);
        float const ix = inv_dir(double(dx)), iy = inv_dir(double(dy)), iz = inv_dir(double(dz));
        int sp = 0;
        int32_t node = 0;
        while (true)
        {
            bvh_node<float> const nd = nodes[node];
            bool const h0 = slab(nd.lo0, nd.hi0, ox, oy, oz, ix, iy, iz);
            bool const h1 = slab(nd.lo1, nd.hi1, ox, oy, oz, ix, iy, iz);
            int32_t next = -1;
            bool have = false;
            int32_t first = 0, count = 0;
            if (h0)
            {
                if (bucket_span<Buckets>(nd.c0, nd.sz & 0xFFu, false, L, first, count))
                    leaves(first, count);
                else
                {
                    next = nd.c0;
                    have = true;
                }
            }
            if (h1)
            {
                if (bucket_span<Buckets>(nd.c1, nd.sz >> 8, true, L, first, count))
                    leaves(first, count);
                else if (have)
                {
                    // both children internal: descend into the one with fewer leaves, push the other (the stack
                    // then never exceeds log2 of the leaf count, < kMaxStack)
                    int32_t const push = nd.big ? nd.c1 : nd.c0;
                    next = nd.big ? nd.c0 : nd.c1;
                    if (sp < kMaxStack)
                        stack[sp++][lane] = push;
                }
                else
                {
                    next = nd.c1;
                    have = true;
                }
            }
            if (have)
            {
                node = next;
            }
            else
            {
                if (sp == 0)
                    break;
                node = stack[--sp][lane];
            }
        }
    }
    for (int32_t b = 0; b < nb; ++b)
        sum += mt_test(rec[brute[b]], px, py, pz, dx, dy, dz);
    return sum;
}

// The same count for a wave of rays at once (packet traversal): the wave walks the LBVH together, carrying the mask
// of lanes whose rays reach the current node; a node's two boxes are tested by those lanes, a leaf's triangle (a
// bucket's triangles) by the lanes whose rays hit its box. Every lane thus tests exactly the triangles the per-lane
// traversal would, with the same arithmetic, so the counts are the same; nodes are read once per wave (scalar loads)
// and the control flow is uniform, which pays when a wave's rays are close and parallel (lattice rows, spatially
// ordered queries). The stack (node, mask) lives in the wave's LDS slots; the larger subtree is pushed, as in
// count_one, so it stays within kMaxStack. packet_walk is the walk itself: it calls leaves(first, count) in each lane
// whose ray hits the box of the leaf (bucket) [first, first + count).
template <bool Buckets, class Leaves>
__device__ __forceinline__ void packet_walk(float ox, float oy, float oz, bool valid, float ix, float iy, float iz,
                                            bvh_node<float> const* __restrict__ nodes, int L, int32_t* stk_node,
                                            uint64_t* stk_mask, int lane, Leaves&& leaves)
{
    uint64_t m = __ballot(valid); // lanes whose rays reach `node` (uniform)
    int32_t node = 0;
    int sp = 0;
    while (m != 0)
    {
        bvh_node<float> const nd = nodes[node];
        bool const act = (m >> lane) & 1ull;
        bool const h0 = act && slab(nd.lo0, nd.hi0, ox, oy, oz, ix, iy, iz);
        bool const h1 = act && slab(nd.lo1, nd.hi1, ox, oy, oz, ix, iy, iz);
        uint64_t const m0 = __ballot(h0), m1 = __ballot(h1);
        int32_t next = -1;
        uint64_t nm = 0;
        int32_t first = 0, count = 0;
        if (m0)
        {
            if (bucket_span<Buckets>(nd.c0, nd.sz & 0xFFu, false, L, first, count))
            {
                if (h0)
                    leaves(first, count);
            }
            else
            {
                next = nd.c0;
                nm = m0;
            }
        }
        if (m1)
        {
            if (bucket_span<Buckets>(nd.c1, nd.sz >> 8, true, L, first, count))
            {
                if (h1)
                    leaves(first, count);
            }
            else if (next >= 0)
            {
                // both children internal: go down the one with fewer leaves, push the other
                if (lane == 0)
                {
                    stk_node[sp] = nd.big ? nd.c1 : nd.c0;
                    stk_mask[sp] = nd.big ? m1 : m0;
                }
                ++sp;
                next = nd.big ? nd.c0 : nd.c1;
                nm = nd.big ? m0 : m1;
            }
            else
            {
                next = nd.c1;
                nm = m1;
            }
        }
        if (next >= 0)
        {
            node = next;
            m = nm;
        }
        else if (sp > 0)
        {
            --sp;
            __builtin_amdgcn_wave_barrier();
            node = __builtin_amdgcn_readfirstlane(stk_node[sp]);
            m = stk_mask[sp];
            __builtin_amdgcn_wave_barrier();
        }
        else
            m = 0;
    }
}

template <bool Buckets, class T>
__device__ __forceinline__ int count_packet(T px, T py, T pz, bool valid, T dx, T dy, T dz,
                                            bvh_node<float> const* __restrict__ nodes,
                                            leaf_rec<T> const* __restrict__ rec, int32_t const* __restrict__ brute,
                                            int32_t nb, int use_bvh, int L, int32_t* stk_node, uint64_t* stk_mask,
                                            int lane)
{
    int sum = 0;
    if (use_bvh)
        packet_walk<Buckets>(float(px), float(py), float(pz), valid, inv_dir(double(dx)), inv_dir(double(dy)),
                             inv_dir(double(dz)), nodes, L, stk_node, stk_mask, lane,
                             [&](int32_t first, int32_t count) __attribute__((always_inline)) {
                                 for (int32_t k = first; k < first + count; ++k)
                                 {
                                     leaf_rec<T> const r = rec[k];
                                     if (r.live)
                                         sum += mt_test(r, px, py, pz, dx, dy, dz);
                                 }
                             });
    if (valid)
        for (int32_t b = 0; b < nb; ++b)
            sum += mt_test(rec[brute[b]], px, py, pz, dx, dy, dz);
    return sum;
}

// Ray order for the packet traversal: key = (query index << 32) | 2D Morton code (16 bits per axis) of the origin
// projected on the plane across the direction (basis u, w), quantized over [-r, r] (r = the largest query norm).
struct ray_key_frame
{
    float u[3], w[3];
    float r, scale; // scale = 65535 / (2 r)
};
constexpr int64_t kSortRaysMin = int64_t(1) << 16; // queries from which count() sorts the rays

__device__ __forceinline__ uint32_t spread2_16(uint32_t v)
{
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}
template <class T>
__global__ void __launch_bounds__(kB) k_ray_keys(T const* __restrict__ pos, int64_t n, ray_key_frame F,
                                                 uint64_t* __restrict__ keys)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n)
        return;
    float const x = float(pos[3 * i]), y = float(pos[3 * i + 1]), z = float(pos[3 * i + 2]);
    float const a = fminf(fmaxf((F.u[0] * x + F.u[1] * y + F.u[2] * z + F.r) * F.scale, 0.0f), 65535.0f); // NaN -> 0
    float const b = fminf(fmaxf((F.w[0] * x + F.w[1] * y + F.w[2] * z + F.r) * F.scale, 0.0f), 65535.0f);
    keys[i] = (uint64_t(i) << 32) | uint64_t(spread2_16(uint32_t(a)) | (spread2_16(uint32_t(b)) << 1));
}
template <class T>
__global__ void __launch_bounds__(kB) k_ray_gather(uint64_t const* __restrict__ sorted, int64_t n,
                                                   T const* __restrict__ pos, T* __restrict__ out)
{
    int64_t const e = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (e >= n)
        return;
    int64_t const i = int64_t(sorted[e] >> 32);
    out[3 * e] = pos[3 * i];
    out[3 * e + 1] = pos[3 * i + 1];
    out[3 * e + 2] = pos[3 * i + 2];
}
__global__ void __launch_bounds__(kB) k_ray_scatter(uint64_t const* __restrict__ sorted, int64_t n,
                                                    int32_t const* __restrict__ cs, int32_t* __restrict__ out)
{
    int64_t const e = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (e < n)
        out[int64_t(sorted[e] >> 32)] = cs[e];
}

// PACKET: the wave walks the tree together (queries in lattice order or close to it); otherwise each lane alone
// (queries in any order: scattered rays share few nodes, and a packet would visit the union of their paths).
template <class T, class Src, bool PACKET>
__global__ void __launch_bounds__(kTB)
    k_count(Src const src, int64_t n, T dx, T dy, T dz, bvh_node<float> const* __restrict__ nodes,
            leaf_rec<T> const* __restrict__ rec, int32_t const* __restrict__ brute, int32_t const* __restrict__ nbrute_p,
            int use_bvh, int32_t* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kTB + threadIdx.x;
    if constexpr (PACKET)
    {
        __shared__ int32_t stk_node[kTB / 64][kMaxStack];
        __shared__ uint64_t stk_mask[kTB / 64][kMaxStack];
        bool const valid = i < n;
        T px = T(0), py = T(0), pz = T(0);
        if (valid)
            query_at(src, i, px, py, pz);
        int const w = int(threadIdx.x >> 6);
        int const c = count_packet<false>(px, py, pz, valid, dx, dy, dz, nodes, rec, brute, *nbrute_p, use_bvh, 1,
                                          stk_node[w], stk_mask[w], int(threadIdx.x & 63));
        if (valid)
            out[i] = c;
    }
    else
    {
        __shared__ int32_t stack[kMaxStack][kTB];
        if (i >= n)
            return;
        T px, py, pz;
        query_at(src, i, px, py, pz);
        out[i] = count_one<false>(px, py, pz, dx, dy, dz, nodes, rec, brute, *nbrute_p, use_bvh, 1, stack,
                                  int(threadIdx.x));
    }
}

// Integer terms of lattice points by rows. When d = +-e_k and the lattice is axis-aligned with one of its axes (L, the
// rows' axis) along e_k, the points of a lattice row differ only in coordinate k, and for such a d the replica's u and v do not depend on
// that coordinate (h_k = d_j e2_l - d_l e2_j = +-0 and d . q = +-q_k, with q_k free of s_k): the points of a row pass
// u and v on the same triangles and differ only in t. A live triangle that the replica counts for a point p has its
// crossing, which shares p's other two coordinates, inside its leaf box (the box holds every point where the replica
// can report a crossing -- the property the traversal relies on). So the triangles any point of a row counts are among
// those whose leaf box spans the row's two coordinates and which pass u and v there: each live triangle is binned into
// those rows (k_row_bin), and each point tests its row's triangles, plus the exhaustive list, with its own coordinates
// -- the same tests as its traversal, so the same count. A row binned more than kRowHits triangles is counted point by
// point.
constexpr int kRowHits = 64;
constexpr int64_t kMaxRows = int64_t(1) << 20; // lattice rows (ny nz) up to which the row counts are used

// Axis-aligned lattice in row form. The rows run along lattice axis L (world axis k); the other two lattice axes P < Q
// step along world axes a and b, and row (iP, iQ) is number iQ nP + iP. Coordinate a of row (iP, .) is sa + ha iP and
// coordinate b of row (., iQ) is sb + hb iQ, the float expressions of query_at with the zero terms dropped (adding a
// zero does not change a nonzero value, and the test's comparisons do not see the sign of a zero).
struct row_frame
{
    int k, a, b;
    int fwd; // d = +e_k
    float sa, ha, sb, hb;
    float ia, ib; // 1 / ha, 1 / hb (estimates only)
    int64_t ny, nz; // nP, nQ
};

// the row of lattice point i, rows along lattice axis L
__device__ __forceinline__ int64_t lattice_row(grid_desc const& g, int L, int64_t i)
{
    int64_t ix, iy, iz;
    if (L == 0)
        return div_recip(i, g.nx, g.inx, ix); // iz ny + iy
    grid_cell(g, i, ix, iy, iz);
    return (L == 1 ? iz : iy) * g.nx + ix;
}

// first index i of [0, n) with lo <= s + h i <= hi in float (none: n); the values are monotone in i, so the indices
// that satisfy the test are contiguous, and the walk from the estimate to their start is exact whatever its error
__device__ __forceinline__ int64_t first_in(float s, float h, float inv_h, int64_t n, float lo, float hi)
{
#pragma clang fp contract(off)
    auto val = [&](int64_t i) { return s + h * float(i); };
    auto before = [&](int64_t i) { return h > 0.0f ? val(i) < lo : val(i) > hi; };
    float const t = ((h > 0.0f ? lo : hi) - s) * inv_h;
    int64_t e = !(t > 0.0f) ? 0 : (t >= float(n - 1) ? n - 1 : int64_t(t));
    while (e > 0 && !before(e - 1))
        --e;
    while (e < n && before(e))
        ++e;
    return e;
}

// The replica's u and v tests (intersector.hh:81-93): true when both pass. The same operations as mt_test.
template <class T>
__device__ __forceinline__ bool mt_uv(leaf_rec<T> const& r, T px, T py, T pz, T dx, T dy, T dz)
{
#pragma clang fp contract(off)
    T const sx = px - r.ax, sy = py - r.ay, sz = pz - r.az;
    T const u = r.inv * (sx * r.hx + sy * r.hy + sz * r.hz);
    if (u < T(0) || u > T(1))
        return false;
    T const qx = sy * r.e1z - sz * r.e1y;
    T const qy = sz * r.e1x - sx * r.e1z;
    T const qz = sx * r.e1y - sy * r.e1x;
    T const v = r.inv * (dx * qx + dy * qy + dz * qz);
    return !(v < T(0) || u + v > T(1));
}

// Bins each live triangle into the rows its leaf box spans where it passes u and v: hits[row][slot] = (triangle, far
// side of its box along d: hi_k for d = +e_k, lo_k for -e_k), nh[row] = the number binned (cleared before; above
// kRowHits the list is incomplete and unused). xk0: coordinate k of the rows' first points (any value gives the same u
// and v).
__global__ void __launch_bounds__(kB)
    k_row_bin(row_frame F, float xk0, int64_t ntri, float dx, float dy, float dz,
              leaf_rec<float> const* __restrict__ rec, float const* __restrict__ lo, float const* __restrict__ hi,
              int2* __restrict__ hits, int32_t* __restrict__ nh)
{
#pragma clang fp contract(off)
    int64_t const t = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (t >= ntri)
        return;
    leaf_rec<float> const r = rec[t];
    if (!r.live)
        return;
    float const la = lo[3 * t + F.a], ha_ = hi[3 * t + F.a], lb = lo[3 * t + F.b], hb_ = hi[3 * t + F.b];
    int64_t const iy0 = first_in(F.sa, F.ha, F.ia, F.ny, la, ha_);
    int64_t const iz0 = first_in(F.sb, F.hb, F.ib, F.nz, lb, hb_);
    for (int64_t iz = iz0; iz < F.nz; ++iz)
    {
        float const zb = F.sb + F.hb * float(iz);
        if (!(lb <= zb && zb <= hb_))
            break;
        for (int64_t iy = iy0; iy < F.ny; ++iy)
        {
            float const ya = F.sa + F.ha * float(iy);
            if (!(la <= ya && ya <= ha_))
                break;
            float const px = F.k == 0 ? xk0 : (F.a == 0 ? ya : zb);
            float const py = F.k == 1 ? xk0 : (F.a == 1 ? ya : zb);
            float const pz = F.k == 2 ? xk0 : (F.a == 2 ? ya : zb);
            if (mt_uv(r, px, py, pz, dx, dy, dz))
            {
                int64_t const row = iz * F.ny + iy;
                int const slot = atomicAdd(&nh[row], 1);
                if (slot < kRowHits)
                    hits[row * kRowHits + slot] = make_int2(int32_t(t), __float_as_int(F.fwd ? hi[3 * t + F.k]
                                                                                              : lo[3 * t + F.k]));
            }
        }
    }
}

// Integer terms from the rows' triangles and the exhaustive list, lane = point: all nx ny nz lattice points (ALL:
// entry e = query e) or m routed entries (positions pts, index-ordered entries li, query index li >> 1). A row's hit
// list holds (triangle, far side of its leaf box along d); a triangle whose box lies wholly behind the point is
// skipped without its record (its test fails: a crossing the test reports lies in the box, ahead of the point). A
// row with an incomplete list takes the packet traversal.
template <bool ALL>
__global__ void __launch_bounds__(kTB)
    k_row_count(grid_desc g, int L, int k, float const* __restrict__ pts, int64_t const* __restrict__ li, int64_t m,
                float dx, float dy, float dz, bvh_node<float> const* __restrict__ nodes,
                leaf_rec<float> const* __restrict__ rec, int32_t const* __restrict__ brute,
                int32_t const* __restrict__ nbrute_p, int2 const* __restrict__ hits, int32_t const* __restrict__ nh,
                int32_t* __restrict__ out)
{
    __shared__ int32_t stk_node[kTB / 64][kMaxStack];
    __shared__ uint64_t stk_mask[kTB / 64][kMaxStack];
    int64_t const e = int64_t(blockIdx.x) * kTB + threadIdx.x;
    bool const valid = e < m;
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    int64_t row = 0;
    int n = 0;
    if (valid)
    {
        if constexpr (ALL)
        {
            query_at(src_grid{g}, e, px, py, pz);
            row = lattice_row(g, L, e);
        }
        else
        {
            px = pts[3 * e];
            py = pts[3 * e + 1];
            pz = pts[3 * e + 2];
            row = lattice_row(g, L, li[e] >> 1);
        }
        n = nh[row];
    }
    int32_t const nb = *nbrute_p;
    int c = 0;
    bool const over = n > kRowHits;
    if (__ballot(valid && over))
        c = count_packet<false>(px, py, pz, valid && over, dx, dy, dz, nodes, rec, brute, nb, 1, 1,
                                stk_node[threadIdx.x >> 6], stk_mask[threadIdx.x >> 6], int(threadIdx.x & 63));
    if (valid && !over)
    {
        float const pk = k == 0 ? px : (k == 1 ? py : pz);
        bool const fwd = (k == 0 ? dx : (k == 1 ? dy : dz)) > 0.0f;
        int2 const* const h = hits + row * kRowHits;
        for (int j = 0; j < n; ++j)
        {
            int2 const hv = h[j];
            float const far = __int_as_float(hv.y);
            if (fwd ? far < pk : far > pk)
                continue;
            c += mt_test(rec[hv.x], px, py, pz, dx, dy, dz);
        }
        for (int32_t b = 0; b < nb; ++b)
            c += mt_test(rec[brute[b]], px, py, pz, dx, dy, dz);
    }
    if (valid)
        out[e] = c;
}

// One prepared mesh of a batch, for the persistent counting kernel.
struct batch_mesh_desc
{
    bvh_node<float> const* nodes;
    leaf_rec<float> const* rec;
    int32_t const* brute;
    int32_t const* nbrute;
    int32_t use_bvh;
    int32_t pad;
    int64_t q_end; // one past the mesh's last query in the concatenated query array
    float kr, kscale; // ray sort keys: projected origins quantized over [-kr, kr], scale = (2^bits - 1) / (2 kr)
};

// Persistent counting kernel over a batch: workgroups take work items (mesh, first query; kTB queries at most, not
// past the mesh's end) from an atomic counter until none are left -- every workgroup ends at the same test. Buckets:
// internal children of at most L leaves are tested as leaf buckets (L is tuned per batch; the counts do not depend
// on it).
template <bool Buckets>
__global__ void __launch_bounds__(kTB)
    k_count_batch(float const* __restrict__ pos, batch_mesh_desc const* __restrict__ meshes,
                  int32_t const* __restrict__ item_mesh, int64_t const* __restrict__ item_q0, int64_t nitems,
                  unsigned long long* __restrict__ counter, float dx, float dy, float dz, int L,
                  int32_t* __restrict__ out)
{
    __shared__ int32_t stack[kMaxStack][kTB];
    __shared__ unsigned long long cur;
    while (true)
    {
        if (threadIdx.x == 0)
            cur = atomicAdd(counter, 1ull);
        __syncthreads();
        unsigned long long const ticket = cur;
        __syncthreads();
        if (ticket >= static_cast<unsigned long long>(nitems)) // unsigned: any counter value past the list ends
            break;
        int64_t const it = int64_t(ticket);
        int32_t const m = item_mesh[it];
        batch_mesh_desc const D = meshes[m];
        int64_t const q = item_q0[it] + threadIdx.x;
        if (q < D.q_end)
            out[q] = count_one<Buckets>(pos[3 * q], pos[3 * q + 1], pos[3 * q + 2], dx, dy, dz, D.nodes, D.rec,
                                        D.brute, D.nbrute ? *D.nbrute : 0, D.use_bvh, L, stack, int(threadIdx.x));
    }
}

// Batched rays in sorted order (full batches of 2^16 queries or more): every query's key is (index << 32) | (mesh <<
// mbits_low) | 2D Morton code of its origin projected across the direction (frame u, w; per-mesh quantization), so a
// sort of the low 32 bits keeps each mesh's queries in its own index range, in ray order; the packet traversal then
// counts them and the counts go back to the callers' order.
struct batch_key_frame
{
    float u[3], w[3];
    int axis_bits; // Morton bits per axis
    int mesh_shift; // 2 * axis_bits
};
__global__ void __launch_bounds__(kTB) k_batch_keys(float const* __restrict__ pos,
                                                   batch_mesh_desc const* __restrict__ meshes,
                                                   int32_t const* __restrict__ item_mesh,
                                                   int64_t const* __restrict__ item_q0, batch_key_frame F,
                                                   uint64_t* __restrict__ keys)
{
    int64_t const it = blockIdx.x;
    int32_t const m = item_mesh[it];
    batch_mesh_desc const D = meshes[m];
    int64_t const q = item_q0[it] + threadIdx.x;
    if (q >= D.q_end)
        return;
    float const x = pos[3 * q], y = pos[3 * q + 1], z = pos[3 * q + 2];
    float const top = float((1u << F.axis_bits) - 1u);
    float const a = fminf(fmaxf((F.u[0] * x + F.u[1] * y + F.u[2] * z + D.kr) * D.kscale, 0.0f), top); // NaN -> 0
    float const b = fminf(fmaxf((F.w[0] * x + F.w[1] * y + F.w[2] * z + D.kr) * D.kscale, 0.0f), top);
    uint32_t const code = spread2_16(uint32_t(a)) | (spread2_16(uint32_t(b)) << 1);
    keys[q] = (uint64_t(q) << 32) | uint64_t((uint32_t(m) << F.mesh_shift) | code);
}

// The persistent counting kernel on sorted queries (the same work items: each mesh keeps its index range), by packet
// traversal: a work item's kTB queries are the block's two waves' packets.
template <bool Buckets>
__global__ void __launch_bounds__(kTB)
    k_count_batch_packet(float const* __restrict__ pos, batch_mesh_desc const* __restrict__ meshes,
                         int32_t const* __restrict__ item_mesh, int64_t const* __restrict__ item_q0, int64_t nitems,
                         unsigned long long* __restrict__ counter, float dx, float dy, float dz, int L,
                         int32_t* __restrict__ out)
{
    __shared__ int32_t stk_node[kTB / 64][kMaxStack];
    __shared__ uint64_t stk_mask[kTB / 64][kMaxStack];
    __shared__ unsigned long long cur;
    int const w = int(threadIdx.x >> 6);
    while (true)
    {
        if (threadIdx.x == 0)
            cur = atomicAdd(counter, 1ull);
        __syncthreads();
        unsigned long long const ticket = cur;
        __syncthreads();
        if (ticket >= static_cast<unsigned long long>(nitems)) // unsigned: any counter value past the list ends
            break;
        int64_t const it = int64_t(ticket);
        int32_t const m = item_mesh[it];
        batch_mesh_desc const D = meshes[m];
        int64_t const q = item_q0[it] + threadIdx.x;
        bool const valid = q < D.q_end;
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        if (valid)
        {
            px = pos[3 * q];
            py = pos[3 * q + 1];
            pz = pos[3 * q + 2];
        }
        int const c = count_packet<Buckets>(px, py, pz, valid, dx, dy, dz, D.nodes, D.rec, D.brute,
                                            D.nbrute ? *D.nbrute : 0, D.use_bvh, L, stk_node[w], stk_mask[w],
                                            int(threadIdx.x & 63));
        if (valid)
            out[q] = c;
    }
}

// Per-direction prepared state, cached on the mesh (rebuilt when the direction changes or the queries reach farther).
template <class T>
struct dir_state
{
    T d[3] = {0, 0, 0};
    double dsum = -1.0;
    leaf_rec<T>* rec = nullptr;
    float* lo = nullptr;
    float* hi = nullptr;
    float* nlo = nullptr;
    float* nhi = nullptr;
    bvh_node<float>* nodes = nullptr;
    int32_t* flags = nullptr;
    int32_t* brute = nullptr;
    int32_t* nbrute = nullptr;
    ~dir_state()
    {
        for (void* p : {(void*)rec, (void*)lo, (void*)hi, (void*)nlo, (void*)nhi, (void*)nodes, (void*)flags,
                        (void*)brute, (void*)nbrute})
            if (p)
                (void)hipFree(p);
    }
};
} // namespace

// ---------------------------------------------------------------------------------------------------------------------
// the mesh handle
// ---------------------------------------------------------------------------------------------------------------------
struct mesh_impl
{
    virtual ~mesh_impl() = default;
    virtual int scalar_bytes() const = 0;
    int64_t ntri = 0;
    int64_t nv = 0;
    lbvh_topology topo;
};

template <class T>
struct mesh_t final : mesh_impl
{
    T* tri = nullptr; // sorted triangle vertices, 9 per triangle
    double vmax = 0.0; // max |v|_2 over the vertices
    std::mutex mu;
    std::shared_ptr<dir_state<T>> cache;

    int scalar_bytes() const override { return int(sizeof(T)); }
    ~mesh_t() override
    {
        cache.reset();
        if (tri)
            (void)hipFree(tri);
        lbvh_release(topo);
    }

    int create(T const* hv, int64_t nverts, int32_t const* hidx, int64_t nidx)
    {
        nv = nverts;
        ntri = nidx / 3;
        for (int64_t i = 0; i < nverts; ++i)
        {
            double const x = double(hv[3 * i]), y = double(hv[3 * i + 1]), z = double(hv[3 * i + 2]);
            vmax = std::max(vmax, std::sqrt(x * x + y * y + z * z));
        }
        if (ntri == 0)
            return ok;
        hipStream_t s = nullptr;
        AGWN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct sguard
        {
            hipStream_t s;
            ~sguard() { (void)hipStreamDestroy(s); }
        } sg{s};
        dbuf dv, di, dc;
        AGWN_TRY(dv.alloc(size_t(std::max<int64_t>(nverts, 1)) * 3 * sizeof(T), s));
        AGWN_TRY(di.alloc(size_t(nidx) * sizeof(int32_t), s));
        AGWN_TRY(dc.alloc(size_t(ntri) * 3 * sizeof(float), s));
        AGWN_HIP(hipMemcpyAsync(dv.p, hv, size_t(nverts) * 3 * sizeof(T), hipMemcpyHostToDevice, s));
        AGWN_HIP(hipMemcpyAsync(di.p, hidx, size_t(nidx) * sizeof(int32_t), hipMemcpyHostToDevice, s));
        unsigned const g = unsigned((ntri + kB - 1) / kB);
        k_centroids<T><<<dim3(g), dim3(kB), 0, s>>>(dv.as<T>(), di.as<int32_t>(), ntri, dc.as<float>());
        AGWN_HIP(hipGetLastError());
        AGWN_TRY(lbvh_build_topology(dc.as<float>(), ntri, topo, s));
        AGWN_HIP(hipMalloc(&tri, size_t(ntri) * 9 * sizeof(T)));
        k_gather_tris<T><<<dim3(g), dim3(kB), 0, s>>>(dv.as<T>(), di.as<int32_t>(), topo.order, ntri, tri);
        AGWN_HIP(hipGetLastError());
        AGWN_HIP(hipStreamSynchronize(s));
        return ok;
    }

    // Returns the prepared state for direction d valid for queries with |p| <= qmax.
    int prepare(T const d[3], double qmax, hipStream_t s, std::shared_ptr<dir_state<T>>& out)
    {
        std::lock_guard<std::mutex> lk(mu);
        double const dsum = qmax + vmax;
        if (cache && cache->d[0] == d[0] && cache->d[1] == d[1] && cache->d[2] == d[2] && dsum <= cache->dsum)
        {
            out = cache;
            return ok;
        }
        auto st = std::make_shared<dir_state<T>>();
        st->d[0] = d[0];
        st->d[1] = d[1];
        st->d[2] = d[2];
        // grow the valid range a little so nearby query sets reuse the state
        st->dsum = dsum * 1.25 + 1e-30;
        size_t const nt = size_t(ntri);
        AGWN_HIP(hipMalloc(&st->rec, nt * sizeof(leaf_rec<T>)));
        AGWN_HIP(hipMalloc(&st->lo, nt * 3 * sizeof(float)));
        AGWN_HIP(hipMalloc(&st->hi, nt * 3 * sizeof(float)));
        AGWN_HIP(hipMalloc(&st->nlo, std::max<size_t>(nt, 2) * 3 * sizeof(float)));
        AGWN_HIP(hipMalloc(&st->nhi, std::max<size_t>(nt, 2) * 3 * sizeof(float)));
        AGWN_HIP(hipMalloc(&st->nodes, std::max<size_t>(nt, 2) * sizeof(bvh_node<float>)));
        AGWN_HIP(hipMalloc(&st->flags, std::max<size_t>(nt, 2) * sizeof(int32_t)));
        AGWN_HIP(hipMalloc(&st->brute, std::max<size_t>(nt, 1) * sizeof(int32_t)));
        AGWN_HIP(hipMalloc(&st->nbrute, sizeof(int32_t)));
        AGWN_HIP(hipMemsetAsync(st->nbrute, 0, sizeof(int32_t), s));
        double const eps = sizeof(T) == 4 ? 0x1p-24 : 0x1p-53;
        double const float_pad = sizeof(T) == 4 ? 0.0 : 4.0 * 0x1p-24 * st->dsum; // float(p) of f64 queries
        double const cond_max = sizeof(T) == 4 ? 0x1p13 : 0x1p40;
        unsigned const g = unsigned((ntri + kB - 1) / kB);
        // a single triangle has no internal node: it is tested exhaustively (unless det == 0: never hit)
        k_prep_leaves<T><<<dim3(g), dim3(kB), 0, s>>>(tri, ntri, d[0], d[1], d[2], st->dsum, eps, float_pad,
                                                      cond_max, ntri == 1 ? 1 : 0, st->rec, st->lo, st->hi, st->brute,
                                                      st->nbrute);
        AGWN_HIP(hipGetLastError());
        if (ntri > 1)
        {
            AGWN_TRY(lbvh_refit<float>(topo, st->lo, st->hi, st->nodes, st->nlo, st->nhi, st->flags, s));
        }
        AGWN_HIP(hipStreamSynchronize(s));
        cache = st;
        out = st;
        return ok;
    }

    // Integer terms of queries in the caller's order. Large calls sort the rays first -- by the 2D Morton code of
    // their origins projected along d, so that a wave's rays run close together -- count them with the packet
    // traversal and scatter the counts back (2M random rays in a 200k-triangle mesh: 2.84 ms per lane in the
    // caller's order, 0.89 ms sorted and in packets). Counts are integers and each ray's triangle tests are the
    // same either way, so the order changes nothing but the time.
    int count(T const* pos, int64_t n, T const d[3], int32_t* out, hipStream_t s)
    {
        if (n == 0)
            return ok;
        if (ntri == 0)
        {
            AGWN_HIP(hipMemsetAsync(out, 0, size_t(n) * sizeof(int32_t), s));
            return ok;
        }
        double qmax = 0.0;
        AGWN_TRY(max_query_norm(pos, n, s, qmax));
        if (n < kSortRaysMin || n >= (int64_t(1) << 32))
            return count_src(src_array<T>{pos}, n, qmax, d, out, s, false);
        double e[3] = {double(d[0]), double(d[1]), double(d[2])};
        double const el = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        AGWN_CHECK(el > 0.0, "zero ray direction");
        for (double& c : e)
            c /= el;
        double const h[3] = {std::fabs(e[0]) < 0.9 ? 1.0 : 0.0, std::fabs(e[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        double u[3] = {e[1] * h[2] - e[2] * h[1], e[2] * h[0] - e[0] * h[2], e[0] * h[1] - e[1] * h[0]};
        double const ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double& c : u)
            c /= ul;
        double const w[3] = {e[1] * u[2] - e[2] * u[1], e[2] * u[0] - e[0] * u[2], e[0] * u[1] - e[1] * u[0]};
        ray_key_frame F{{float(u[0]), float(u[1]), float(u[2])}, {float(w[0]), float(w[1]), float(w[2])},
                        float(qmax), qmax > 0.0 ? float(65535.0 / (2.0 * qmax)) : 0.0f};
        dbuf keys, keys2, ps, cs;
        AGWN_TRY(keys.alloc(size_t(n) * sizeof(uint64_t), s));
        AGWN_TRY(keys2.alloc(size_t(n) * sizeof(uint64_t), s));
        AGWN_TRY(ps.alloc(size_t(n) * 3 * sizeof(T), s));
        AGWN_TRY(cs.alloc(size_t(n) * sizeof(int32_t), s));
        unsigned const g = unsigned((n + kB - 1) / kB);
        k_ray_keys<T><<<dim3(g), dim3(kB), 0, s>>>(pos, n, F, keys.as<uint64_t>());
        AGWN_HIP(hipGetLastError());
        uint64_t* sorted = nullptr;
        AGWN_TRY(radix_sort_u64(keys.as<uint64_t>(), keys2.as<uint64_t>(), n, 32, s, &sorted));
        k_ray_gather<T><<<dim3(g), dim3(kB), 0, s>>>(sorted, n, pos, ps.as<T>());
        AGWN_HIP(hipGetLastError());
        AGWN_TRY(count_src(src_array<T>{ps.as<T>()}, n, qmax, d, cs.as<int32_t>(), s, true));
        k_ray_scatter<<<dim3(g), dim3(kB), 0, s>>>(sorted, n, cs.as<int32_t>(), out);
        AGWN_HIP(hipGetLastError());
        return ok;
    }

    static int max_query_norm(T const* pos, int64_t n, hipStream_t s, double& qmax)
    {
        dbuf qm;
        AGWN_TRY(qm.alloc(sizeof(unsigned long long), s));
        AGWN_HIP(hipMemsetAsync(qm.p, 0, sizeof(unsigned long long), s));
        unsigned const g = unsigned(std::min<int64_t>((n + kB - 1) / kB, 2048));
        k_max_norm<T><<<dim3(g), dim3(kB), 0, s>>>(pos, n, qm.as<unsigned long long>());
        AGWN_HIP(hipGetLastError());
        unsigned long long bits = 0;
        AGWN_HIP(hipMemcpyAsync(&bits, qm.p, sizeof(bits), hipMemcpyDeviceToHost, s));
        AGWN_HIP(hipStreamSynchronize(s));
        std::memcpy(&qmax, &bits, sizeof(qmax));
        return ok;
    }

    // counts of the queries of `src`, all of norm <= qmax
    template <class Src>
    int count_src(Src const& src, int64_t n, double qmax, T const d[3], int32_t* out, hipStream_t s, bool packet)
    {
        if (n == 0)
            return ok;
        if (ntri == 0)
        {
            AGWN_HIP(hipMemsetAsync(out, 0, size_t(n) * sizeof(int32_t), s));
            return ok;
        }
        std::shared_ptr<dir_state<T>> st;
        AGWN_TRY(prepare(d, qmax, s, st));
        int64_t const blocks = (n + kTB - 1) / kTB;
        AGWN_CHECK(blocks < (int64_t(1) << 31), "too many queries (%lld)", (long long)n);
        if (packet)
            k_count<T, Src, true><<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(
                src, n, d[0], d[1], d[2], st->nodes, st->rec, st->brute, st->nbrute, ntri > 1 ? 1 : 0, out);
        else
            k_count<T, Src, false><<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(
                src, n, d[0], d[1], d[2], st->nodes, st->rec, st->brute, st->nbrute, ntri > 1 ? 1 : 0, out);
        AGWN_HIP(hipGetLastError());
        // keep the state alive until the kernel has consumed it (it may be replaced by another thread's call)
        AGWN_HIP(hipStreamSynchronize(s));
        return ok;
    }

    // Integer terms by rows (k_row_bin) when d = +-e_k and lattice g is axis-aligned with one axis along e_k; *done =
    // false (nothing launched) otherwise. li null: all m = nx ny nz lattice points, in the lattice's order; else m
    // routed points (positions pts; entries li in index order, query index li >> 1). All of norm <= qmax. Temporaries
    // come from ws, or are allocated on s when ws is null.
    int count_rows(grid_desc const& g, int64_t const* li, float const* pts, int64_t m, double qmax, float const d[3],
                   int32_t* out, hipStream_t s, arena* ws, bool& done)
    {
        static_assert(std::is_same_v<T, float>, "row counts are FP32");
        done = false;
        if (m == 0 || ntri < 2)
            return ok;
        // the axis of each lattice step's one nonzero component (-1: none or several)
        auto axis = [](float const* v) {
            int ax = -1, nz = 0;
            for (int c = 0; c < 3; ++c)
                if (v[c] != 0.0f)
                {
                    ax = c;
                    ++nz;
                }
            return nz == 1 ? ax : -1;
        };
        float const* step[3] = {g.dx, g.dy, g.dz};
        int64_t const cnt[3] = {g.nx, g.ny, g.nz};
        int const ax[3] = {axis(g.dx), axis(g.dy), axis(g.dz)};
        int const k = axis(d);
        if (k < 0 || std::fabs(d[k]) != 1.0f || ax[0] < 0 || ax[1] < 0 || ax[2] < 0 || ax[0] == ax[1] ||
            ax[0] == ax[2] || ax[1] == ax[2])
            return ok;
        int const L = ax[0] == k ? 0 : (ax[1] == k ? 1 : 2); // the lattice axis along the ray
        int const P = L == 0 ? 1 : 0, Q = L == 2 ? 1 : 2;
        int64_t const rows = cnt[P] * cnt[Q];
        if (rows > kMaxRows)
            return ok;
        int const a = ax[P], b = ax[Q];
        row_frame const F{k, a, b, d[k] > 0.0f ? 1 : 0, g.start[a], step[P][a], g.start[b], step[Q][b],
                          1.0f / step[P][a], 1.0f / step[Q][b], cnt[P], cnt[Q]};
        std::shared_ptr<dir_state<T>> st;
        AGWN_TRY(prepare(d, qmax, s, st));
        int32_t* nh = nullptr;
        int2* hits = nullptr;
        dbuf bn, bh;
        if (ws)
        {
            AGWN_TRY(ws->take_n(size_t(rows), &nh));
            AGWN_TRY(ws->take_n(size_t(rows) * kRowHits, &hits));
        }
        else
        {
            AGWN_TRY(bn.alloc(size_t(rows) * sizeof(int32_t), s));
            AGWN_TRY(bh.alloc(size_t(rows) * kRowHits * sizeof(int2), s));
            nh = bn.as<int32_t>();
            hits = bh.as<int2>();
        }
        AGWN_HIP(hipMemsetAsync(nh, 0, size_t(rows) * sizeof(int32_t), s));
        k_row_bin<<<dim3(unsigned((ntri + kB - 1) / kB)), dim3(kB), 0, s>>>(F, g.start[k], ntri, d[0], d[1], d[2],
                                                                         st->rec, st->lo, st->hi, hits, nh);
        AGWN_HIP(hipGetLastError());
        int64_t const blocks = (m + kTB - 1) / kTB;
        AGWN_CHECK(blocks < (int64_t(1) << 31), "too many queries (%lld)", (long long)m);
        if (li)
            k_row_count<false><<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(
                g, L, k, pts, li, m, d[0], d[1], d[2], st->nodes, st->rec, st->brute, st->nbrute, hits, nh, out);
        else
            k_row_count<true><<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(g, L, k, nullptr, nullptr, m, d[0], d[1], d[2],
                                                                           st->nodes, st->rec, st->brute, st->nbrute,
                                                                           hits, nh, out);
        AGWN_HIP(hipGetLastError());
        AGWN_HIP(hipStreamSynchronize(s)); // the state outlives its kernels
        done = true;
        return ok;
    }
};
} // namespace agwn

struct agwn_mesh
{
    std::unique_ptr<agwn::mesh_impl> impl;
    mutable agwn::plan_cache boundary_plan; // FP32: the last boundary's chain plan (the boundary comes per call)
    // point labelling's device workspace (agwn_label_surrogate_dev), reused between calls; the calls use it one at a
    // time under label_mu and finish their device work before they return
    mutable std::mutex label_mu;
    mutable std::unique_ptr<agwn::arena> label_ws;
    mutable int label_ws_dev = -1;
};

namespace agwn
{
namespace
{
struct call_stream2
{
    hipStream_t s = nullptr;
    ~call_stream2()
    {
        if (s)
            (void)hipStreamDestroy(s);
    }
};

template <class F>
int guarded2(F&& f)
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
}

template <class T>
int mesh_create(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, agwn_mesh** out)
{
    AGWN_CHECK(out != nullptr, "out is null");
    *out = nullptr;
    AGWN_CHECK(nv >= 0 && nidx >= 0, "negative count");
    AGWN_CHECK(nidx % 3 == 0, "indices.size() = %lld is not a multiple of 3", (long long)nidx);
    AGWN_CHECK(nidx == 0 || (verts && idx), "null mesh arrays");
    for (int64_t i = 0; i < nidx; ++i)
        AGWN_CHECK(idx[i] >= 0 && idx[i] < nv, "index %lld = %d out of range [0, %lld)", (long long)i, idx[i],
                   (long long)nv);
    AGWN_TRY(ensure_device());
    auto m = std::make_unique<mesh_t<T>>();
    AGWN_TRY(m->create(verts, nv, idx, nidx));
    auto* h = new agwn_mesh;
    h->impl = std::move(m);
    *out = h;
    return ok;
}

template <class T>
mesh_t<T>* typed(agwn_mesh const* m)
{
    if (!m || !m->impl || m->impl->scalar_bytes() != int(sizeof(T)))
        return nullptr;
    return static_cast<mesh_t<T>*>(m->impl.get());
}

template <class T>
int int_dev(agwn_mesh const* m, T const* pos, int64_t n, T const* dir, int32_t* out, hipStream_t s)
{
    mesh_t<T>* mt = typed<T>(m);
    AGWN_CHECK(mt != nullptr, "mesh handle is null or of the other scalar type");
    AGWN_CHECK(n >= 0, "negative count");
    AGWN_CHECK(dir != nullptr, "direction is null");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    T const d[3] = {dir[0], dir[1], dir[2]};
    return mt->count(pos, n, d, out, s);
}

template <class T>
int full_dev(agwn_mesh const* m, T const* pos, int64_t n, T const* x0, T const* seg7, int64_t nseg, T* out,
             hipStream_t s)
{
    mesh_t<T>* mt = typed<T>(m);
    AGWN_CHECK(mt != nullptr, "mesh handle is null or of the other scalar type");
    AGWN_CHECK(n >= 0 && nseg >= 0, "negative count");
    AGWN_CHECK(x0 != nullptr, "x0 is null");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    if (n == 0)
        return ok;
    dbuf cnt;
    AGWN_TRY(cnt.alloc(size_t(n) * sizeof(int32_t), s));
    T const d[3] = {x0[0], x0[1], x0[2]};
    AGWN_TRY(mt->count(pos, n, d, cnt.as<int32_t>(), s));
    double const x[3] = {double(x0[0]), double(x0[1]), double(x0[2])};
    if constexpr (sizeof(T) == 4)
    {
        std::shared_ptr<frac_plan const> pl;
        AGWN_TRY(cached_plan(m->boundary_plan, seg7, nseg, s, pl));
        return frac_f32_planned(src_array<float>{pos}, n, *pl, seg7, nseg, x, cnt.as<int32_t>(), out, s);
    }
    else
        return frac_f64(pos, n, seg7, nseg, x, cnt.as<int32_t>(), out, s, true);
}

template <class T>
int int_host(agwn_mesh const* m, T const* pos, int64_t n, T const* dir, int32_t* out)
{
    AGWN_CHECK(typed<T>(m) != nullptr, "mesh handle is null or of the other scalar type");
    AGWN_CHECK(n >= 0, "negative count");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    if (n == 0)
        return ok;
    call_stream2 cs;
    AGWN_HIP(hipStreamCreateWithFlags(&cs.s, hipStreamNonBlocking));
    dbuf dp, dout;
    AGWN_TRY(dp.alloc(size_t(n) * 3 * sizeof(T), cs.s));
    AGWN_TRY(dout.alloc(size_t(n) * sizeof(int32_t), cs.s));
    AGWN_HIP(hipMemcpyAsync(dp.p, pos, size_t(n) * 3 * sizeof(T), hipMemcpyHostToDevice, cs.s));
    AGWN_TRY(int_dev<T>(m, dp.as<T>(), n, dir, dout.as<int32_t>(), cs.s));
    AGWN_HIP(hipMemcpyAsync(out, dout.p, size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost, cs.s));
    AGWN_HIP(hipStreamSynchronize(cs.s));
    return ok;
}

template <class T>
int full_host(agwn_mesh const* m, T const* pos, int64_t n, T const* x0, T const* seg7, int64_t nseg, T* out)
{
    AGWN_CHECK(typed<T>(m) != nullptr, "mesh handle is null or of the other scalar type");
    AGWN_CHECK(n >= 0 && nseg >= 0, "negative count");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    if (n == 0)
        return ok;
    call_stream2 cs;
    AGWN_HIP(hipStreamCreateWithFlags(&cs.s, hipStreamNonBlocking));
    dbuf dp, ds, dout;
    AGWN_TRY(dp.alloc(size_t(n) * 3 * sizeof(T), cs.s));
    AGWN_TRY(ds.alloc(size_t(std::max<int64_t>(nseg, 1)) * 7 * sizeof(T), cs.s));
    AGWN_TRY(dout.alloc(size_t(n) * sizeof(T), cs.s));
    AGWN_HIP(hipMemcpyAsync(dp.p, pos, size_t(n) * 3 * sizeof(T), hipMemcpyHostToDevice, cs.s));
    if (nseg > 0)
        AGWN_HIP(hipMemcpyAsync(ds.p, seg7, size_t(nseg) * 7 * sizeof(T), hipMemcpyHostToDevice, cs.s));
    AGWN_TRY(full_dev<T>(m, dp.as<T>(), n, x0, nseg ? ds.as<T>() : nullptr, nseg, dout.as<T>(), cs.s));
    AGWN_HIP(hipMemcpyAsync(out, dout.p, size_t(n) * sizeof(T), hipMemcpyDeviceToHost, cs.s));
    AGWN_HIP(hipStreamSynchronize(cs.s));
    return ok;
}
} // namespace
} // namespace agwn

using namespace agwn;

extern "C" {
int agwn_mesh_create_f32(float const* v, int64_t nv, int32_t const* idx, int64_t nidx, agwn_mesh** out)
{
    return guarded2([&] { return mesh_create(v, nv, idx, nidx, out); });
}
int agwn_mesh_create_f64(double const* v, int64_t nv, int32_t const* idx, int64_t nidx, agwn_mesh** out)
{
    return guarded2([&] { return mesh_create(v, nv, idx, nidx, out); });
}
int agwn_mesh_destroy(agwn_mesh* m)
{
    clear_error();
    delete m;
    return ok;
}
int agwn_mesh_info(agwn_mesh const* m, int64_t* ntri, int64_t* nnodes, int* scalar_bytes)
{
    clear_error();
    AGWN_CHECK(m && m->impl, "mesh handle is null");
    if (ntri)
        *ntri = m->impl->ntri;
    if (nnodes)
        *nnodes = m->impl->ntri > 1 ? m->impl->ntri - 1 : 0;
    if (scalar_bytes)
        *scalar_bytes = m->impl->scalar_bytes();
    return ok;
}
int agwn_int_batch_f32(agwn_mesh const* m, float const* pos, int64_t n, float const* dir, int32_t* out)
{
    return guarded2([&] { return int_host(m, pos, n, dir, out); });
}
int agwn_int_batch_f64(agwn_mesh const* m, double const* pos, int64_t n, double const* dir, int32_t* out)
{
    return guarded2([&] { return int_host(m, pos, n, dir, out); });
}
int agwn_int_batch_f32_dev(agwn_mesh const* m, float const* pos, int64_t n, float const* dir, int32_t* out,
                           agwn_stream s)
{
    return guarded2([&] { return int_dev(m, pos, n, dir, out, hipStream_t(s)); });
}
int agwn_int_batch_f64_dev(agwn_mesh const* m, double const* pos, int64_t n, double const* dir, int32_t* out,
                           agwn_stream s)
{
    return guarded2([&] { return int_dev(m, pos, n, dir, out, hipStream_t(s)); });
}
int agwn_full_batch_f32(agwn_mesh const* m, float const* pos, int64_t n, float const* x0, float const* seg7,
                        int64_t nseg, float* out)
{
    return guarded2([&] { return full_host(m, pos, n, x0, seg7, nseg, out); });
}
int agwn_full_batch_f64(agwn_mesh const* m, double const* pos, int64_t n, double const* x0, double const* seg7,
                        int64_t nseg, double* out)
{
    return guarded2([&] { return full_host(m, pos, n, x0, seg7, nseg, out); });
}
int agwn_full_batch_f32_dev(agwn_mesh const* m, float const* pos, int64_t n, float const* x0, float const* seg7,
                            int64_t nseg, float* out, agwn_stream s)
{
    return guarded2([&] { return full_dev(m, pos, n, x0, seg7, nseg, out, hipStream_t(s)); });
}
int agwn_full_batch_f64_dev(agwn_mesh const* m, double const* pos, int64_t n, double const* x0, double const* seg7,
                            int64_t nseg, double* out, agwn_stream s)
{
    return guarded2([&] { return full_dev(m, pos, n, x0, seg7, nseg, out, hipStream_t(s)); });
}
} // extern "C"

// =====================================================================================================================
// Dense-grid evaluator: OpenGLMeshGwn::set_scene / eval_grid (gwn_mesh_opengl.hh:426-643) as HIP kernels that
// generate the lattice points themselves and write the full GWN straight to HBM in the caller's
// ((iz*ny)+iy)*nx+ix layout -- no readback, no host un-permute. The ray / pole direction is chosen as the reference
// does: the axis with the fewest cells (ties to the lower axis), x0 = normalize(that axis' step) (:521-550); the full
// GWN does not depend on that choice. Each cell's value is eval_gwnr_mesh_batch of its lattice point with that x0:
// integer term bit-exact with NaiveIntersector<float>, fractional as the FP32 fractional kernel.
// =====================================================================================================================
struct agwn_grid
{
    std::mutex mu;
    std::unique_ptr<agwn::mesh_t<float>> mesh;
    float* seg7 = nullptr;
    int64_t nseg = 0;
    std::unique_ptr<agwn::frac_plan> plan; // the boundary's chain plan, built once per scene
    // Routed-query list of the labelling calls (8 bytes per query), kept between calls and grown when a call needs
    // more: allocating and freeing it per call costs about 3 ms at 512^3 (the free of a 1 GiB block). The calls use it
    // one at a time under mu, and each finishes its device work before it returns.
    void* label_list = nullptr;
    size_t label_list_bytes = 0;
    int label_list_dev = -1;
    // the labelling calls' other device temporaries (routed points, sort keys and scratch, undecided lists, ...),
    // reused the same way: per-call allocation and release cost about 3 ms per call at 512^3, in the releases
    std::unique_ptr<agwn::arena> label_ws;
    ~agwn_grid()
    {
        if (seg7)
            (void)hipFree(seg7);
        if (label_list)
            (void)hipFree(label_list);
    }
};

namespace agwn
{
namespace
{
int grid_set_scene(agwn_grid* g, float const* verts, int64_t nv, int32_t const* idx, int64_t nidx, float const* seg7,
                   int64_t nseg)
{
    AGWN_CHECK(g != nullptr, "grid handle is null");
    AGWN_CHECK(nv >= 0 && nidx >= 0 && nseg >= 0, "negative count");
    AGWN_CHECK(nidx % 3 == 0, "indices.size() = %lld is not a multiple of 3", (long long)nidx);
    AGWN_CHECK(nidx == 0 || (verts && idx), "null mesh arrays");
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    for (int64_t i = 0; i < nidx; ++i)
        AGWN_CHECK(idx[i] >= 0 && idx[i] < nv, "index %lld = %d out of range [0, %lld)", (long long)i, idx[i],
                   (long long)nv);
    AGWN_TRY(ensure_device());
    auto m = std::make_unique<mesh_t<float>>();
    AGWN_TRY(m->create(verts, nv, idx, nidx));
    auto plan = std::make_unique<frac_plan>();
    AGWN_TRY(frac_plan_build(seg7, nseg, *plan, nullptr));
    float* d = nullptr;
    if (nseg > 0)
    {
        AGWN_HIP(hipMalloc(&d, size_t(nseg) * 7 * sizeof(float)));
        hipError_t const e = hipMemcpy(d, seg7, size_t(nseg) * 7 * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess)
        {
            (void)hipFree(d);
            AGWN_HIP(e);
        }
    }
    std::lock_guard<std::mutex> lk(g->mu);
    g->mesh = std::move(m);
    if (g->seg7)
        (void)hipFree(g->seg7);
    g->seg7 = d;
    g->nseg = nseg;
    g->plan = std::move(plan);
    return ok;
}

// The lattice, the integration / pole axis and the query-norm bound of one eval_grid call.
struct grid_call
{
    grid_desc gd;
    int64_t n;
    float x0f[3];
    double x0d[3];
    double qmax;
};

int grid_setup(float const* start, float const* dx, float const* dy, float const* dz, int nx, int ny, int nz,
               grid_call& c)
{
#pragma clang fp contract(off)
    AGWN_CHECK(start && dx && dy && dz, "null grid vector");
    AGWN_CHECK(nx > 0 && ny > 0 && nz > 0, "grid counts must be positive (%d, %d, %d)", nx, ny, nz);
    for (int k = 0; k < 3; ++k)
    {
        c.gd.start[k] = start[k];
        c.gd.dx[k] = dx[k];
        c.gd.dy[k] = dy[k];
        c.gd.dz[k] = dz[k];
    }
    AGWN_CHECK(int64_t(nx) * ny <= (int64_t(1) << 52) / nz, "grid of %d x %d x %d cells is too large", nx, ny, nz);
    c.gd.nx = nx;
    c.gd.ny = ny;
    c.gd.nz = nz;
    c.gd.inx = 1.0 / double(nx);
    c.gd.iny = 1.0 / double(ny);
    c.n = int64_t(nx) * ny * nz;
    // integration / pole axis: fewest cells, ties to the lower axis (gwn_mesh_opengl.hh:524-531)
    float const* dirs[3] = {dx, dy, dz};
    int const counts[3] = {nx, ny, nz};
    int ia = 0;
    if (counts[1] < counts[ia])
        ia = 1;
    if (counts[2] < counts[ia])
        ia = 2;
    float const* a = dirs[ia];
    float const len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    AGWN_CHECK(len > 0.0f, "the integration axis step is zero");
    for (int k = 0; k < 3; ++k)
    {
        c.x0f[k] = a[k] / len;
        c.x0d[k] = double(c.x0f[k]);
    }
    // largest query norm over the lattice corners (the lattice is the convex hull of its corners)
    double qmax = 0.0;
    for (int cc = 0; cc < 8; ++cc)
    {
        double const fx = (cc & 1) ? double(nx - 1) : 0.0, fy = (cc & 2) ? double(ny - 1) : 0.0,
                     fz = (cc & 4) ? double(nz - 1) : 0.0;
        double n2 = 0.0;
        for (int k = 0; k < 3; ++k)
        {
            double const v = double(start[k]) + double(dx[k]) * fx + double(dy[k]) * fy + double(dz[k]) * fz;
            n2 += v * v;
        }
        qmax = std::max(qmax, std::sqrt(n2));
    }
    c.qmax = qmax * (1.0 + 1e-5) + 1e-30;
    return ok;
}

int grid_eval_dev(agwn_grid* g, float const* start, float const* dx, float const* dy, float const* dz, int nx, int ny,
                  int nz, float* out, hipStream_t s)
{
    AGWN_CHECK(g != nullptr, "grid handle is null");
    AGWN_CHECK(out != nullptr, "out is null");
    grid_call c;
    AGWN_TRY(grid_setup(start, dx, dy, dz, nx, ny, nz, c));
    std::lock_guard<std::mutex> lk(g->mu);
    AGWN_CHECK(g->mesh != nullptr, "set_scene was not called");
    dbuf cnt;
    AGWN_TRY(cnt.alloc(size_t(c.n) * sizeof(int32_t), s));
    bool rows = false;
    AGWN_TRY(g->mesh->count_rows(c.gd, nullptr, nullptr, c.n, c.qmax, c.x0f, cnt.as<int32_t>(), s, nullptr, rows));
    if (!rows)
        AGWN_TRY(g->mesh->count_src(src_grid{c.gd}, c.n, c.qmax, c.x0f, cnt.as<int32_t>(), s, true));
    return frac_f32_planned(src_grid{c.gd}, c.n, *g->plan, g->seg7, g->nseg, c.x0d, cnt.as<int32_t>(), out, s);
}

// Gather / scatter of the points the far-field decision left open (list = indices into the routed arrays).
__global__ void __launch_bounds__(kB) k_gather_undecided(float const* __restrict__ pts, int32_t const* __restrict__ cnt,
                                                         int64_t const* __restrict__ list, int64_t n,
                                                         float* __restrict__ opts, int32_t* __restrict__ ocnt)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= n)
        return;
    int64_t const i = list[k];
    opts[3 * k] = pts[3 * i];
    opts[3 * k + 1] = pts[3 * i + 1];
    opts[3 * k + 2] = pts[3 * i + 2];
    ocnt[k] = cnt[i];
}
__global__ void __launch_bounds__(kB) k_scatter_values(int64_t const* __restrict__ list, int64_t n,
                                                       float const* __restrict__ v, float* __restrict__ out)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k < n)
        out[list[k]] = v[k];
}

// Routed entries, positions and integer terms from index order to the decision's order: out[e] = in[perm[e]].
__global__ void __launch_bounds__(kB) k_permute_routed(uint64_t const* __restrict__ perm, int64_t m,
                                                       int64_t const* __restrict__ li, float const* __restrict__ pi,
                                                       int32_t const* __restrict__ ci, int64_t* __restrict__ lo,
                                                       float* __restrict__ po, int32_t* __restrict__ co)
{
    int64_t const e = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (e >= m)
        return;
    int64_t const k = int64_t(perm[e]);
    lo[e] = li[k];
    po[3 * e] = pi[3 * k];
    po[3 * e + 1] = pi[3 * k + 1];
    po[3 * e + 2] = pi[3 * k + 2];
    co[e] = ci[k];
}

// Values of m routed points (positions pts, integer terms cnt) on the correct side of 0.5: decision at the given
// level (0: kDecideBeta, 1: kDecideBetaTight); the points it leaves open go to the next level, after the last to the
// full boundary sum.
int label_exact_levels(float const* pts, int32_t const* cnt, int64_t m, frac_plan const& pl, float const* seg7,
                       int64_t nseg, double const x0d[3], int level, float* w, hipStream_t s, arena& ws)
{
    if (m == 0)
        return ok;
    if (level == 2) // labels only: small launches may split the stream
        return frac_f32_planned(src_array<float>{pts}, m, pl, seg7, nseg, x0d, cnt, w, s, true, &ws);
    int64_t* und = nullptr;
    AGWN_TRY(ws.take_n(size_t(m), &und));
    int64_t nu = 0;
    AGWN_TRY(frac_decide(pts, m, cnt, pl, x0d, level == 0 ? kDecideBeta : kDecideBetaTight, w, und, nu, s, &ws));
    if (nu == 0)
        return ok;
    float *up = nullptr, *uw = nullptr;
    int32_t* uc = nullptr;
    AGWN_TRY(ws.take_n(size_t(nu) * 3, &up));
    AGWN_TRY(ws.take_n(size_t(nu), &uc));
    AGWN_TRY(ws.take_n(size_t(nu), &uw));
    unsigned const gb = unsigned((nu + kB - 1) / kB);
    k_gather_undecided<<<dim3(gb), dim3(kB), 0, s>>>(pts, cnt, und, nu, up, uc);
    AGWN_HIP(hipGetLastError());
    AGWN_TRY(label_exact_levels(up, uc, nu, pl, seg7, nseg, x0d, level + 1, uw, s, ws));
    k_scatter_values<<<dim3(gb), dim3(kB), 0, s>>>(und, nu, uw, w);
    AGWN_HIP(hipGetLastError());
    return ok;
}

// Integer terms of the m routed points (positions pts, entries li in index order): by lattice rows where the lattice
// and the direction allow it (mesh_t::count_rows), else by the packet traversal over the points.
template <class Src>
int count_routed(mesh_t<float>& mesh, Src const& src, int64_t const* li, float const* pts, int64_t m, double qmax,
                 float const x0f[3], int32_t* c, hipStream_t s, arena& ws)
{
    if constexpr (std::is_same_v<Src, src_grid>)
    {
        bool done = false;
        AGWN_TRY(mesh.count_rows(src.g, li, pts, m, qmax, x0f, c, s, &ws, done));
        if (done)
            return ok;
    }
    return mesh.count_src(src_array<float>{pts}, m, qmax, x0f, c, s, true);
}

// Surrogate labelling with the exact FP32 path as fallback (see mlp_route): network labels for every query, then the
// mask, band and audit points are gathered and labelled exactly: their integer terms (the same kernel, x0 and query
// coordinates as the plain evaluation), then the far-field decision (frac_decide) where the boundary is made of
// closed loops, and the full FP32 boundary sum for the points it leaves open (or for all, otherwise). A decided
// point's value is an estimate of its GWN farther from 0.5 than the far-field remainder bound (rigorous) plus a margin
// for the FP32 evaluations' own rounding (empirical: 1e-4, 26 times the largest FP32 error measured), so it lies on the
// same side of 0.5 as the full evaluation.
template <class Src>
int label_surrogate(mesh_t<float>& mesh, agwn_mlp const* mlp, Src const& src, int64_t n, double qmax,
                    float const x0f[3], float const* seg7, int64_t nseg, frac_plan const* plan, float delta,
                    int64_t audit_every, uint64_t audit_seed, int8_t* labels, int64_t* stats, hipStream_t s,
                    arena& ws, int64_t* list_ws = nullptr)
{
    AGWN_CHECK(mlp != nullptr, "mlp handle is null");
    AGWN_CHECK(delta >= 0.0f, "delta must be >= 0 (+inf routes every point)"); // rejects NaN
    AGWN_CHECK(audit_every >= 0 && audit_every <= int64_t(UINT32_MAX), "audit_every out of range");
    AGWN_CHECK(n == 0 || labels != nullptr, "labels is null");
    AGWN_CHECK(nseg == 0 || seg7 != nullptr, "boundary is null with nseg=%lld", (long long)nseg);
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    if (n > 0)
    {
        // every device temporary comes from the caller's workspace, which the previous call (using it under the same
        // lock) left with its device work finished; the routed list has room for every query (list_ws: n entries of
        // the caller's, else from the workspace)
        ws.reset();
        unsigned long long* dc = nullptr;
        AGWN_TRY(ws.take_n(5, &dc));
        int64_t* list = list_ws;
        if (!list)
            AGWN_TRY(ws.take_n(size_t(n), &list));
        float* lat = nullptr; // layer-0 table of an axis-separable lattice
        if constexpr (std::is_same_v<Src, src_grid>)
        {
            int64_t const tf = mlp_lattice_table_floats(src.g);
            if (tf > 0)
                AGWN_TRY(ws.take_n(size_t(tf), &lat));
        }
        AGWN_HIP(hipMemsetAsync(dc, 0, sizeof(cnt), s));
        AGWN_TRY(mlp_route(mlp, src, n, delta, uint32_t(audit_every), uint32_t(audit_seed ^ (audit_seed >> 32)), labels,
                           list, dc, s, lat));
        AGWN_HIP(hipMemcpyAsync(cnt, dc, sizeof(cnt), hipMemcpyDeviceToHost, s));
        AGWN_HIP(hipStreamSynchronize(s));
        int64_t const m = int64_t(cnt[3]);
        if (m > 0)
        {
            // the list was appended tile by tile in completion order. The integer terms take it in index order
            // (on a lattice, a wave's rays then run along a few neighbouring rows: the packet traversal shares most
            // of their nodes); the decision takes it in Morton order (a wave's queries then need the same loops).
            uint64_t *tmp = nullptr, *keys = nullptr, *keys2 = nullptr;
            uint32_t* sort_ws = nullptr;
            AGWN_TRY(ws.take_n(size_t(m), &tmp));
            AGWN_TRY(ws.take_n(size_t(m), &keys));
            AGWN_TRY(ws.take_n(size_t(m), &keys2));
            AGWN_TRY(ws.take_n(size_t(radix_sort_words(m)), &sort_ws));
            int64_t const* Li = nullptr;
            uint64_t const* perm = nullptr;
            AGWN_TRY(route_order(mlp, src, n, list, m, tmp, keys, keys2, sort_ws, s, &Li, &perm));
            double const x0d[3] = {double(x0f[0]), double(x0f[1]), double(x0f[2])};
            float *pts = nullptr, *w = nullptr;
            int32_t* c = nullptr;
            AGWN_TRY(ws.take_n(size_t(m) * 3, &pts));
            AGWN_TRY(ws.take_n(size_t(m), &c));
            AGWN_TRY(ws.take_n(size_t(m), &w));
            int64_t const* L = Li;
            if (perm)
            {
                float* pts_i = nullptr;
                int32_t* c_i = nullptr;
                int64_t* lm = nullptr;
                AGWN_TRY(ws.take_n(size_t(m) * 3, &pts_i));
                AGWN_TRY(ws.take_n(size_t(m), &c_i));
                AGWN_TRY(ws.take_n(size_t(m), &lm));
                AGWN_TRY(gather_routed(src, Li, m, pts_i, s));
                AGWN_TRY(count_routed(mesh, src, Li, pts_i, m, qmax, x0f, c_i, s, ws));
                k_permute_routed<<<dim3(unsigned((m + kB - 1) / kB)), dim3(kB), 0, s>>>(perm, m, Li, pts_i, c_i,
                                                                                        lm, pts, c);
                AGWN_HIP(hipGetLastError());
                L = lm;
            }
            else
            {
                AGWN_TRY(gather_routed(src, Li, m, pts, s));
                AGWN_TRY(count_routed(mesh, src, Li, pts, m, qmax, x0f, c, s, ws));
            }
            frac_plan local;
            frac_plan const* pl = plan;
            if (!pl)
            {
                std::vector<float> h(size_t(nseg) * 7);
                if (nseg > 0)
                {
                    AGWN_HIP(hipMemcpyAsync(h.data(), seg7, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
                    AGWN_HIP(hipStreamSynchronize(s));
                }
                AGWN_TRY(frac_plan_build(h.data(), nseg, local, s));
                pl = &local;
            }
            if (pl->chain && pl->loops_ok)
            {
                // far-field decision, then a tighter one over what it leaves; only the points neither decides take
                // the full boundary sum. Each level gathers its open points (positions, counts) and scatters the
                // values back.
                AGWN_TRY(label_exact_levels(pts, c, m, *pl, seg7, nseg, x0d, 0, w, s, ws));
            }
            else
                AGWN_TRY(frac_f32_planned(src_array<float>{pts}, m, *pl, seg7, nseg, x0d, c, w, s, false, &ws));
            AGWN_TRY(route_apply(L, m, w, labels, dc, s));
            AGWN_HIP(hipMemcpyAsync(cnt, dc, sizeof(cnt), hipMemcpyDeviceToHost, s));
            AGWN_HIP(hipStreamSynchronize(s));
        }
    }
    if (stats)
        for (int k = 0; k < 5; ++k)
            stats[k] = int64_t(cnt[k]);
    return ok;
}

int grid_label_dev(agwn_grid* g, agwn_mlp const* mlp, float const* start, float const* dx, float const* dy,
                   float const* dz, int nx, int ny, int nz, float delta, int64_t audit_every, uint64_t audit_seed,
                   int8_t* labels, int64_t* stats, hipStream_t s)
{
    AGWN_CHECK(g != nullptr, "grid handle is null");
    grid_call c;
    AGWN_TRY(grid_setup(start, dx, dy, dz, nx, ny, nz, c));
    std::lock_guard<std::mutex> lk(g->mu);
    AGWN_CHECK(g->mesh != nullptr, "set_scene was not called");
    size_t const need = size_t(c.n) * sizeof(int64_t);
    int dev = 0;
    AGWN_HIP(hipGetDevice(&dev));
    if (g->label_list && (g->label_list_bytes < need || g->label_list_dev != dev))
    {
        AGWN_HIP(hipFree(g->label_list));
        g->label_list = nullptr;
        g->label_list_bytes = 0;
    }
    if (g->label_list_dev != dev)
        g->label_ws.reset(); // blocks of another device
    if (!g->label_ws)
        g->label_ws = std::make_unique<arena>();
    if (!g->label_list && need > 0)
    {
        AGWN_HIP(hipMalloc(&g->label_list, need));
        g->label_list_bytes = need;
        g->label_list_dev = dev;
    }
    g->label_list_dev = dev;
    int const rc = label_surrogate(*g->mesh, mlp, src_grid{c.gd}, c.n, c.qmax, c.x0f, g->seg7, g->nseg,
                                   g->plan.get(), delta, audit_every, audit_seed, labels, stats, s, *g->label_ws,
                                   static_cast<int64_t*>(g->label_list));
    if (rc != ok)
        (void)hipStreamSynchronize(s); // a failed call may have left work queued on the list: drain it
    return rc;
}

int grid_eval_host(agwn_grid* g, float const* start, float const* dx, float const* dy, float const* dz, int nx, int ny,
                   int nz, float* out)
{
    AGWN_CHECK(nx > 0 && ny > 0 && nz > 0, "grid counts must be positive (%d, %d, %d)", nx, ny, nz);
    AGWN_CHECK(out != nullptr, "out is null");
    size_t const n = size_t(nx) * size_t(ny) * size_t(nz);
    call_stream2 cs;
    AGWN_HIP(hipStreamCreateWithFlags(&cs.s, hipStreamNonBlocking));
    dbuf d;
    AGWN_TRY(d.alloc(n * sizeof(float), cs.s));
    AGWN_TRY(grid_eval_dev(g, start, dx, dy, dz, nx, ny, nz, d.as<float>(), cs.s));
    AGWN_HIP(hipMemcpyAsync(out, d.p, n * sizeof(float), hipMemcpyDeviceToHost, cs.s));
    AGWN_HIP(hipStreamSynchronize(cs.s));
    return ok;
}
} // namespace
} // namespace agwn

extern "C" {
int agwn_grid_create(agwn_grid** out)
{
    clear_error();
    AGWN_CHECK(out != nullptr, "out is null");
    *out = new agwn_grid;
    return ok;
}
int agwn_grid_destroy(agwn_grid* g)
{
    clear_error();
    delete g;
    return ok;
}
int agwn_grid_set_scene(agwn_grid* g, float const* verts, int64_t nv, int32_t const* idx, int64_t nidx,
                        float const* seg7, int64_t nseg)
{
    return guarded2([&] { return grid_set_scene(g, verts, nv, idx, nidx, seg7, nseg); });
}
int agwn_grid_eval(agwn_grid* g, float const* start, float const* dx, float const* dy, float const* dz, int nx, int ny,
                   int nz, float* out)
{
    return guarded2([&] { return grid_eval_host(g, start, dx, dy, dz, nx, ny, nz, out); });
}
int agwn_grid_eval_dev(agwn_grid* g, float const* start, float const* dx, float const* dy, float const* dz, int nx,
                       int ny, int nz, float* out, agwn_stream s)
{
    return guarded2([&] { return grid_eval_dev(g, start, dx, dy, dz, nx, ny, nz, out, hipStream_t(s)); });
}
int agwn_grid_label_surrogate_dev(agwn_grid* g, agwn_mlp const* mlp, float const* start, float const* dx,
                                  float const* dy, float const* dz, int nx, int ny, int nz, float delta,
                                  int64_t audit_every, uint64_t audit_seed, int8_t* labels, int64_t* stats,
                                  agwn_stream s)
{
    return guarded2([&] {
        return grid_label_dev(g, mlp, start, dx, dy, dz, nx, ny, nz, delta, audit_every, audit_seed, labels, stats,
                              hipStream_t(s));
    });
}
int agwn_mlp_predict_grid_dev(agwn_mlp const* mlp, float const* start, float const* dx, float const* dy,
                              float const* dz, int nx, int ny, int nz, float* out, agwn_stream s)
{
    return guarded2([&]() -> int {
        grid_call c;
        AGWN_TRY(grid_setup(start, dx, dy, dz, nx, ny, nz, c));
        return mlp_predict_grid_dev(mlp, c.gd, out, hipStream_t(s));
    });
}
int agwn_label_surrogate_dev(agwn_mesh const* m, agwn_mlp const* mlp, float const* pos, int64_t n, float const* x0,
                             float const* seg7, int64_t nseg, float delta, int64_t audit_every, uint64_t audit_seed,
                             int8_t* labels, int64_t* stats, agwn_stream s)
{
    return guarded2([&]() -> int {
        AGWN_CHECK(m != nullptr && m->impl != nullptr, "mesh handle is null");
        AGWN_CHECK(m->impl->scalar_bytes() == 4, "surrogate labelling needs an f32 mesh");
        AGWN_CHECK(n >= 0, "negative count");
        AGWN_CHECK(n == 0 || pos != nullptr, "positions are null");
        AGWN_CHECK(x0 != nullptr, "x0 is null");
        auto& mesh = static_cast<mesh_t<float>&>(*m->impl);
        hipStream_t const st = hipStream_t(s);
        double qmax = 0.0;
        if (n > 0)
            AGWN_TRY(mesh_t<float>::max_query_norm(pos, n, st, qmax));
        std::shared_ptr<frac_plan const> pl;
        if (n > 0)
            AGWN_TRY(cached_plan(m->boundary_plan, seg7, nseg, st, pl));
        std::lock_guard<std::mutex> lk(m->label_mu);
        int dev = 0;
        AGWN_HIP(hipGetDevice(&dev));
        if (!m->label_ws || m->label_ws_dev != dev)
        {
            m->label_ws = std::make_unique<arena>();
            m->label_ws_dev = dev;
        }
        int const rc = label_surrogate(mesh, mlp, src_array<float>{pos}, n, qmax, x0, seg7, nseg, pl.get(), delta,
                                       audit_every, audit_seed, labels, stats, st, *m->label_ws);
        if (rc != ok)
            (void)hipStreamSynchronize(st); // a failed call may have left work queued on the workspace: drain it
        return rc;
    });
}
} // extern "C"

// =====================================================================================================================
// Batched evaluation of many meshes with uneven query counts: the full GWN of every query of every mesh, in two
// persistent launches (the integer term, then the fractional term plus the counts) captured once in a HIP graph and
// replayed while the inputs stay the same. Each mesh keeps its own LBVH and chain-ordered boundary stream; work items
// (mesh, query tile) are ordered longest mesh first and taken from atomic counters. An autotuner picks the fractional
// kernel's shape (queries per lane x block x stream points per LDS stage), timing every shape on a sample of the
// meshes, and the counting traversal's leaf-bucket size, timing every size on the whole batch. Per query the arithmetic
// is exactly the single-mesh path's (k_count, k_frac_f32_chain with x0 shared by the batch) for every choice.
// =====================================================================================================================
struct agwn_batch
{
    std::mutex mu;
    int64_t k = 0;
    std::vector<std::unique_ptr<agwn::mesh_t<float>>> meshes;
    std::vector<int64_t> pt_off; // k + 1 offsets into the point stream
    std::vector<int64_t> nunit;  // unit boundary segments per mesh
    std::vector<int> sexp;       // per mesh: boundary_scale_exp of its boundary (the FP32 coordinate scale)
    std::vector<float4> far;     // per mesh: its far-query test
    agwn::dbuf stream;           // the concatenated point streams, caller frame (persistent)
    agwn::dbuf pexp;             // int16 per stream point: its mesh's sexp
    agwn::dbuf zero;             // an int32 zero: the exhaustive-test count of meshes without a BVH state
    hipStream_t bs = nullptr;    // the batch's own stream (graph capture and replay)
    int variant = -1;            // tuned fractional kernel shape (index into the shape table)
    int leaf = -1;               // tuned leaf-bucket size (index into kLeafSizes)
    double tune_ms[32] = {};     // the tuner's time per fractional shape
    double leaf_ms[8] = {};      // the tuner's time per leaf-bucket size
    int ncu = 0;
    // the cached launch: inputs it was built for, its buffers and graph
    float const* g_pos = nullptr;
    float* g_out = nullptr;
    float g_x0[3] = {0, 0, 0};
    std::vector<int64_t> g_qoff;
    int g_variant = -1, g_leaf = -1;
    std::vector<std::shared_ptr<agwn::dir_state<float>>> states;
    agwn::dbuf cdesc, fdesc, citem_m, citem_q, fitem_m, fitem_q, counters, counts, p32;
    int64_t ncitems = 0, nfitems = 0;
    // rays in sorted order for the integer term (full batches of kSortRaysMin queries or more)
    bool sorted = false;
    int64_t g_sorted_n = 0; // queries of the sorted launch
    agwn::batch_key_frame kf{};
    agwn::dbuf skeys, skeys2, sws, spos, scnt;
    hipGraphExec_t exec = nullptr;
    ~agwn_batch()
    {
        if (exec)
            (void)hipGraphExecDestroy(exec);
        if (bs)
        {
            (void)hipStreamSynchronize(bs);
            (void)hipStreamDestroy(bs);
        }
    }
};

namespace agwn
{
// leaf-bucket sizes offered to the autotuner (1: plain leaves)
constexpr int kLeafSizes[] = {1, 2, 4, 8, 16};
constexpr int kNumLeafSizes = int(sizeof(kLeafSizes) / sizeof(kLeafSizes[0]));
static_assert(kLeafSizes[kNumLeafSizes - 1] <= kMaxBucket, "bucket size");
static_assert(kNumLeafSizes == AGWN_BATCH_LEAF_SIZES, "the header's count of leaf-bucket sizes");
namespace
{
__global__ void __launch_bounds__(kB) k_zero_u64(unsigned long long* p, int n)
{
    if (int(threadIdx.x) < n)
        p[threadIdx.x] = 0ull;
}

// per-mesh largest query norm (as float bits: non-negative floats order like their bit patterns). Each lane takes
// kNormPerLane queries strided by the block width (coalesced) and keeps a running maximum for the mesh it is in,
// flushing it with one atomic when the queries move on to a later mesh; at the end the lanes of a wave that all
// ended in the same mesh (the common case: a mesh's queries are contiguous) reduce first and issue one atomic.
constexpr int kNormPerLane = 16;
__global__ void __launch_bounds__(kB) k_seg_max_norm(float const* __restrict__ pos, int64_t n,
                                                     int64_t const* __restrict__ qoff, int64_t k,
                                                     unsigned int* __restrict__ out)
{
    int64_t const i0 = int64_t(blockIdx.x) * (kB * kNormPerLane) + threadIdx.x;
    int64_t m = -1;   // mesh of the running maximum
    unsigned int r = 0; // running maximum; 0 = none (every norm below is >= 1e-30)
    if (i0 < n)
    {
        // mesh of query i0: the last offset <= i0 (binary search over the k + 1 offsets)
        int64_t lo = 0, hi = k;
        while (hi - lo > 1)
        {
            int64_t const mid = (lo + hi) / 2;
            if (qoff[mid] <= i0)
                lo = mid;
            else
                hi = mid;
        }
        m = lo;
        for (int j = 0; j < kNormPerLane; ++j)
        {
            int64_t const i = i0 + int64_t(j) * kB;
            if (i >= n)
                break;
            while (qoff[m + 1] <= i) // a later mesh (qoff[k] = n > i ends this; meshes without queries are passed)
            {
                if (r)
                    atomicMax(out + m, r);
                r = 0;
                ++m;
            }
            double const x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
            float const nrm = float(sqrt(x * x + y * y + z * z)) * 1.00001f + 1e-30f;
            r = max(r, __float_as_uint(nrm));
        }
    }
    // every lane of the (full) block reaches this point: the wave-wide vote and shuffles see all 64 lanes
    int const mi = int(m);
    int const m0 = __shfl(mi, 0);
    if (__all(mi == m0))
    {
        for (int off = 32; off > 0; off >>= 1)
            r = max(r, __shfl_xor(r, off));
        if ((threadIdx.x & 63) == 0 && m0 >= 0 && r)
            atomicMax(out + m0, r);
    }
    else if (m >= 0 && r)
        atomicMax(out + m, r);
}

int batch_create(int64_t k, float const* verts, int64_t const* voff, int32_t const* idx, int64_t const* ioff,
                 float const* seg7, int64_t const* soff, agwn_batch** out)
{
    AGWN_CHECK(out != nullptr, "out is null");
    *out = nullptr;
    AGWN_CHECK(k > 0 && voff && ioff && soff, "a batch needs k > 0 meshes and their offset arrays");
    AGWN_CHECK(k < (int64_t(1) << 31), "a batch holds at most 2^31 - 1 meshes (%lld)", (long long)k);
    AGWN_CHECK(voff[0] == 0 && ioff[0] == 0 && soff[0] == 0, "offsets must start at 0");
    for (int64_t m = 0; m < k; ++m)
    {
        AGWN_CHECK(voff[m + 1] >= voff[m] && ioff[m + 1] >= ioff[m] && soff[m + 1] >= soff[m],
                   "offsets must be non-decreasing (mesh %lld)", (long long)m);
        AGWN_CHECK((ioff[m + 1] - ioff[m]) % 3 == 0, "mesh %lld: index count not a multiple of 3", (long long)m);
    }
    AGWN_CHECK(voff[k] == 0 || verts, "verts is null");
    AGWN_CHECK(ioff[k] == 0 || idx, "idx is null");
    AGWN_CHECK(soff[k] == 0 || seg7, "seg7 is null");
    AGWN_TRY(ensure_device());
    auto b = std::make_unique<agwn_batch>();
    b->k = k;
    b->pt_off.assign(size_t(k + 1), 0);
    b->nunit.assign(size_t(k), 0);
    b->sexp.assign(size_t(k), 0);
    b->far.assign(size_t(k), make_float4(0.0f, 0.0f, 0.0f, HUGE_VALF));
    std::vector<float4> all;
    std::vector<int16_t> allexp;
    for (int64_t m = 0; m < k; ++m)
    {
        int64_t const nv = voff[m + 1] - voff[m];
        int64_t const ni = ioff[m + 1] - ioff[m];
        int32_t const* im = idx + ioff[m];
        for (int64_t i = 0; i < ni; ++i)
            AGWN_CHECK(im[i] >= 0 && im[i] < nv, "mesh %lld: index %lld = %d out of range [0, %lld)", (long long)m,
                       (long long)i, im[i], (long long)nv);
        auto mesh = std::make_unique<mesh_t<float>>();
        AGWN_TRY(mesh->create(verts + 3 * voff[m], nv, im, ni));
        b->meshes.push_back(std::move(mesh));
        std::vector<float4> st;
        int64_t nu = 0;
        AGWN_CHECK(chain_stream(seg7 + 7 * soff[m], soff[m + 1] - soff[m], st, nu),
                   "mesh %lld: batched evaluation needs integer boundary weights (|w| <= 16)", (long long)m);
        float4 far{};
        int const e = boundary_scale_exp(seg7 + 7 * soff[m], soff[m + 1] - soff[m], &far);
        all.insert(all.end(), st.begin(), st.end());
        allexp.insert(allexp.end(), st.size(), int16_t(e));
        b->pt_off[size_t(m + 1)] = int64_t(all.size());
        b->nunit[size_t(m)] = nu;
        b->sexp[size_t(m)] = e;
        b->far[size_t(m)] = far;
    }
    AGWN_HIP(hipStreamCreateWithFlags(&b->bs, hipStreamNonBlocking));
    AGWN_TRY(b->stream.alloc(std::max<size_t>(all.size(), 1) * sizeof(float4), nullptr));
    AGWN_TRY(b->pexp.alloc(std::max<size_t>(allexp.size(), 1) * sizeof(int16_t), nullptr));
    if (!all.empty())
    {
        AGWN_HIP(hipMemcpy(b->stream.p, all.data(), all.size() * sizeof(float4), hipMemcpyHostToDevice));
        AGWN_HIP(hipMemcpy(b->pexp.p, allexp.data(), allexp.size() * sizeof(int16_t), hipMemcpyHostToDevice));
    }
    AGWN_TRY(b->zero.alloc(sizeof(int32_t), nullptr));
    AGWN_HIP(hipMemset(b->zero.p, 0, sizeof(int32_t)));
    int dev = 0;
    AGWN_HIP(hipGetDevice(&dev));
    AGWN_HIP(hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    AGWN_HIP(hipDeviceSynchronize());
    *out = b.release();
    return ok;
}

// Prepares the per-direction state of every mesh and the work lists for `variant` (meshes listed in `which`, all
// when empty) into the batch's buffers. Synchronous.
// per-mesh bound on the query norms (synchronous)
int batch_qmax(agwn_batch& B, float const* pos, int64_t const* qoff, std::vector<float>& qmax)
{
    hipStream_t const s = B.bs;
    int64_t const k = B.k;
    int64_t const n = qoff[k];
    qmax.assign(size_t(k), 0.0f);
    if (n == 0)
        return ok;
    dbuf dq, mx;
    AGWN_TRY(dq.alloc(size_t(k + 1) * sizeof(int64_t), s));
    AGWN_TRY(mx.alloc(size_t(k) * sizeof(unsigned), s));
    AGWN_HIP(hipMemcpyAsync(dq.p, qoff, size_t(k + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    AGWN_HIP(hipMemsetAsync(mx.p, 0, size_t(k) * sizeof(unsigned), s));
    int64_t const per_block = int64_t(kB) * kNormPerLane;
    k_seg_max_norm<<<dim3(unsigned((n + per_block - 1) / per_block)), dim3(kB), 0, s>>>(pos, n, dq.as<int64_t>(), k,
                                                                                         mx.as<unsigned>());
    AGWN_HIP(hipGetLastError());
    AGWN_HIP(hipMemcpyAsync(qmax.data(), mx.p, size_t(k) * sizeof(float), hipMemcpyDeviceToHost, s));
    AGWN_HIP(hipStreamSynchronize(s));
    return ok;
}

int batch_prepare(agwn_batch& B, float const* pos, int64_t const* qoff, float const x0f[3], int variant,
                  std::vector<int64_t> const& which)
{
    hipStream_t const s = B.bs;
    int64_t const k = B.k;
    int64_t const n = qoff[k];
    std::vector<float> qmax;
    AGWN_TRY(batch_qmax(B, pos, qoff, qmax));
    std::vector<char> use(size_t(k), which.empty() ? 1 : 0);
    for (int64_t m : which)
        use[size_t(m)] = 1;
    std::vector<batch_mesh_desc> cd(static_cast<size_t>(k));
    std::vector<batch_frac_desc> fd(static_cast<size_t>(k));
    B.states.assign(size_t(k), nullptr);
    for (int64_t m = 0; m < k; ++m)
    {
        mesh_t<float>& M = *B.meshes[size_t(m)];
        batch_mesh_desc d{};
        d.nbrute = B.zero.as<int32_t>();
        d.q_end = qoff[m + 1];
        if (use[size_t(m)] && qoff[m + 1] > qoff[m] && M.ntri > 0)
        {
            std::shared_ptr<dir_state<float>> st;
            AGWN_TRY(M.prepare(x0f, double(qmax[size_t(m)]), s, st));
            d.nodes = st->nodes;
            d.rec = st->rec;
            d.brute = st->brute;
            d.nbrute = st->nbrute;
            d.use_bvh = M.ntri > 1 ? 1 : 0;
            B.states[size_t(m)] = st;
        }
        cd[size_t(m)] = d;
        fd[size_t(m)] = batch_frac_desc{B.pt_off[size_t(m)], B.pt_off[size_t(m + 1)] - B.pt_off[size_t(m)], qoff[m + 1],
                                        std::ldexp(1.0, -B.sexp[size_t(m)]), B.far[size_t(m)]};
    }
    // work items, longest first: counting by triangle count, the fractional term by stream length
    int fq = 0, fb = 0, fs = 0;
    frac_batch_shape(variant, fq, fb, fs);
    std::vector<int64_t> order(static_cast<size_t>(k));
    for (int64_t m = 0; m < k; ++m)
        order[size_t(m)] = m;
    std::vector<int32_t> cim, fim;
    std::vector<int64_t> ciq, fiq;
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return B.meshes[size_t(a)]->ntri > B.meshes[size_t(b)]->ntri;
    });
    for (int64_t m : order) // meshes without triangles too: their counts must be written (as 0)
        if (use[size_t(m)])
            for (int64_t q = qoff[m]; q < qoff[m + 1]; q += kTB)
            {
                cim.push_back(int32_t(m));
                ciq.push_back(q);
            }
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        return B.pt_off[size_t(a) + 1] - B.pt_off[size_t(a)] > B.pt_off[size_t(b) + 1] - B.pt_off[size_t(b)];
    });
    for (int64_t m : order)
        if (use[size_t(m)])
            for (int64_t q = qoff[m]; q < qoff[m + 1]; q += int64_t(fq) * fb)
            {
                fim.push_back(int32_t(m));
                fiq.push_back(q);
            }
    B.ncitems = int64_t(cim.size());
    B.nfitems = int64_t(fim.size());
    auto up = [&](dbuf& d, void const* h, size_t bytes) -> int {
        AGWN_TRY(d.alloc(std::max<size_t>(bytes, 16), nullptr));
        if (bytes)
            AGWN_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
        return ok;
    };
    AGWN_TRY(up(B.cdesc, cd.data(), cd.size() * sizeof(batch_mesh_desc)));
    AGWN_TRY(up(B.fdesc, fd.data(), fd.size() * sizeof(batch_frac_desc)));
    AGWN_TRY(up(B.citem_m, cim.data(), cim.size() * sizeof(int32_t)));
    AGWN_TRY(up(B.citem_q, ciq.data(), ciq.size() * sizeof(int64_t)));
    AGWN_TRY(up(B.fitem_m, fim.data(), fim.size() * sizeof(int32_t)));
    AGWN_TRY(up(B.fitem_q, fiq.data(), fiq.size() * sizeof(int64_t)));
    AGWN_TRY(B.counters.alloc(2 * sizeof(unsigned long long), nullptr));
    AGWN_HIP(hipMemsetAsync(B.counters.p, 0, 2 * sizeof(unsigned long long), s)); // prepared outside any capture
    AGWN_TRY(B.counts.alloc(size_t(std::max<int64_t>(n, 1)) * sizeof(int32_t), nullptr));
    // sorted rays for the integer term: whole batches only (the tuner's samples leave other queries without keys),
    // with room for the mesh index and at least 4 Morton bits per axis in the 32-bit sort key
    int mbits = 0;
    while ((int64_t(1) << mbits) < k)
        ++mbits;
    int const axis_bits = std::min(15, (32 - mbits) / 2); // mesh_shift = 2 axis_bits <= 30: a defined shift
    B.sorted = which.empty() && n >= kSortRaysMin && n < (int64_t(1) << 32) && axis_bits >= 4;
    B.g_sorted_n = n;
    if (B.sorted)
    {
        double const e0 = double(x0f[0]), e1 = double(x0f[1]), e2 = double(x0f[2]);
        double const el = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
        AGWN_CHECK(el > 0.0, "zero ray direction");
        double const e[3] = {e0 / el, e1 / el, e2 / el};
        double const h[3] = {std::fabs(e[0]) < 0.9 ? 1.0 : 0.0, std::fabs(e[0]) < 0.9 ? 0.0 : 1.0, 0.0};
        double u[3] = {e[1] * h[2] - e[2] * h[1], e[2] * h[0] - e[0] * h[2], e[0] * h[1] - e[1] * h[0]};
        double const ul = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double& c : u)
            c /= ul;
        double const w[3] = {e[1] * u[2] - e[2] * u[1], e[2] * u[0] - e[0] * u[2], e[0] * u[1] - e[1] * u[0]};
        for (int a = 0; a < 3; ++a)
        {
            B.kf.u[a] = float(u[a]);
            B.kf.w[a] = float(w[a]);
        }
        B.kf.axis_bits = axis_bits;
        B.kf.mesh_shift = 2 * axis_bits;
        for (int64_t m = 0; m < k; ++m)
        {
            double const r = qmax[size_t(m)] > 0.0f ? double(qmax[size_t(m)]) : 1.0;
            cd[size_t(m)].kr = float(r);
            cd[size_t(m)].kscale = float(double((1u << axis_bits) - 1u) / (2.0 * r));
        }
        AGWN_TRY(B.skeys.alloc(size_t(n) * sizeof(uint64_t), nullptr));
        AGWN_TRY(B.skeys2.alloc(size_t(n) * sizeof(uint64_t), nullptr));
        AGWN_TRY(B.sws.alloc(size_t(radix_sort_words(n)) * sizeof(uint32_t), nullptr));
        AGWN_TRY(B.spos.alloc(size_t(n) * 3 * sizeof(float), nullptr));
        AGWN_TRY(B.scnt.alloc(size_t(n) * sizeof(int32_t), nullptr));
        AGWN_HIP(hipMemcpyAsync(B.cdesc.p, cd.data(), cd.size() * sizeof(batch_mesh_desc), hipMemcpyHostToDevice, s));
    }
    int64_t const npts = B.pt_off[size_t(k)];
    AGWN_TRY(B.p32.alloc(size_t(std::max<int64_t>(npts, 1)) * 2 * sizeof(float4), nullptr));
    AGWN_HIP(hipStreamSynchronize(s));
    return ok;
}

// The batch's launches for the prepared lists, by phase: kPhaseCount -- counter reset and the integer term (sorted
// rays: keys, sort, gather, packet traversal, scatter; else the per-lane traversal) with leaf buckets of
// kLeafSizes[leaf]; kPhaseFrac -- stream rotation, counter reset, fractional term + counts with shape `variant`. The
// fractional kernel reads the counts, so a whole evaluation runs both, in this order. kWarm: only the phases'
// persistent kernels, one workgroup with no work items each (it exits at once; loads the kernel's code).
constexpr int kPhaseCount = 1, kPhaseFrac = 2, kWarm = 4;

int batch_launches(agwn_batch& B, float const* pos, float const x0f[3], int variant, int leaf, float* out,
                   int phases = kPhaseCount | kPhaseFrac)
{
    hipStream_t const s = B.bs;
    double const x0d[3] = {double(x0f[0]), double(x0f[1]), double(x0f[2])};
    int const L = kLeafSizes[leaf];
    if (phases & kWarm)
    {
        unsigned long long* const ctr = B.counters.as<unsigned long long>();
        k_zero_u64<<<dim3(1), dim3(kB), 0, s>>>(ctr, 2); // the warm-up kernels take a ticket each
        AGWN_HIP(hipGetLastError());
        if (phases & kPhaseCount)
        {
            auto kc = L > 1 ? k_count_batch_packet<true> : k_count_batch_packet<false>;
            auto ku = L > 1 ? k_count_batch<true> : k_count_batch<false>;
            (B.sorted ? kc : ku)<<<dim3(1), dim3(kTB), 0, s>>>(pos, B.cdesc.as<batch_mesh_desc>(),
                                                               B.citem_m.as<int32_t>(), B.citem_q.as<int64_t>(), 0,
                                                               ctr, x0f[0], x0f[1], x0f[2], L,
                                                               B.counts.as<int32_t>());
            AGWN_HIP(hipGetLastError());
        }
        if (phases & kPhaseFrac)
            AGWN_TRY(frac_batch_launch(variant, 1, pos, B.fdesc.as<batch_frac_desc>(), B.fitem_m.as<int32_t>(),
                                       B.fitem_q.as<int64_t>(), 0, ctr + 1, B.p32.as<float4>(), x0d,
                                       B.counts.as<int32_t>(), out, s));
        return ok;
    }
    if ((phases & kPhaseCount) && B.ncitems > 0)
    {
        k_zero_u64<<<dim3(1), dim3(kB), 0, s>>>(B.counters.as<unsigned long long>(), 1);
        AGWN_HIP(hipGetLastError());
        if (B.sorted)
        {
            int64_t const n = B.g_sorted_n;
            k_batch_keys<<<dim3(unsigned(B.ncitems)), dim3(kTB), 0, s>>>(pos, B.cdesc.as<batch_mesh_desc>(),
                                                                        B.citem_m.as<int32_t>(),
                                                                        B.citem_q.as<int64_t>(), B.kf,
                                                                        B.skeys.as<uint64_t>());
            AGWN_HIP(hipGetLastError());
            uint64_t* sorted = nullptr;
            AGWN_TRY(radix_sort_u64_ws(B.skeys.as<uint64_t>(), B.skeys2.as<uint64_t>(), n, 32, B.sws.as<uint32_t>(),
                                       s, &sorted));
            unsigned const g = unsigned((n + kB - 1) / kB);
            k_ray_gather<float><<<dim3(g), dim3(kB), 0, s>>>(sorted, n, pos, B.spos.as<float>());
            AGWN_HIP(hipGetLastError());
            int const blocks = int(std::min<int64_t>(B.ncitems, int64_t(B.ncu) * 16));
            auto kc = L > 1 ? k_count_batch_packet<true> : k_count_batch_packet<false>;
            kc<<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(B.spos.as<float>(), B.cdesc.as<batch_mesh_desc>(),
                                                            B.citem_m.as<int32_t>(), B.citem_q.as<int64_t>(),
                                                            B.ncitems, B.counters.as<unsigned long long>(), x0f[0],
                                                            x0f[1], x0f[2], L, B.scnt.as<int32_t>());
            AGWN_HIP(hipGetLastError());
            k_ray_scatter<<<dim3(g), dim3(kB), 0, s>>>(sorted, n, B.scnt.as<int32_t>(), B.counts.as<int32_t>());
            AGWN_HIP(hipGetLastError());
        }
        else
        {
            // 10 traversal blocks per CU: the 14 KiB LDS stacks allow 11; 5 waves per SIMD hide the traversal's loads
            int const blocks = int(std::min<int64_t>(B.ncitems, int64_t(B.ncu) * 10));
            auto kc = L > 1 ? k_count_batch<true> : k_count_batch<false>;
            kc<<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(pos, B.cdesc.as<batch_mesh_desc>(), B.citem_m.as<int32_t>(),
                                                            B.citem_q.as<int64_t>(), B.ncitems,
                                                            B.counters.as<unsigned long long>(), x0f[0], x0f[1],
                                                            x0f[2], L, B.counts.as<int32_t>());
            AGWN_HIP(hipGetLastError());
        }
    }
    if ((phases & kPhaseFrac) && B.nfitems > 0)
    {
        int64_t const npts = B.pt_off[size_t(B.k)];
        AGWN_TRY(frac_stream_rotate(B.stream.as<float4 const>(), npts, x0d, B.pexp.as<int16_t const>(),
                                    B.p32.as<float4>(), s));
        k_zero_u64<<<dim3(1), dim3(kB), 0, s>>>(B.counters.as<unsigned long long>() + 1, 1);
        AGWN_HIP(hipGetLastError());
        int const blocks = int(std::min<int64_t>(B.nfitems, int64_t(B.ncu) * frac_batch_blocks_per_cu(variant)));
        AGWN_TRY(frac_batch_launch(variant, blocks, pos, B.fdesc.as<batch_frac_desc>(), B.fitem_m.as<int32_t>(),
                                   B.fitem_q.as<int64_t>(), B.nfitems, B.counters.as<unsigned long long>() + 1,
                                   B.p32.as<float4>(), x0d, B.counts.as<int32_t>(), out, s));
    }
    return ok;
}

// Autotuner. The two launches are independent in cost, so each is tuned alone: every fractional kernel shape is timed
// on every 8th mesh (the same work for each shape), then every leaf-bucket size on the whole batch's integer term
// (the path its evaluations take: sorted packets or per-lane rays). Each kernel is launched once with no work items
// before it is timed (its code is then loaded). The fastest of each is kept.
int batch_tune(agwn_batch& B, float const* pos, int64_t const* qoff, float const x0f[3], float* out)
{
    std::vector<int64_t> sample;
    for (int64_t m = 0; m < B.k; m += 8)
        sample.push_back(m);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    AGWN_HIP(hipEventCreate(&e0));
    AGWN_HIP(hipEventCreate(&e1));
    struct eguard
    {
        hipEvent_t a, b;
        ~eguard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    auto timed = [&](int variant, int leaf, int phase, double& ms_out) -> int {
        AGWN_TRY(batch_launches(B, pos, x0f, variant, leaf, out, phase | kWarm));
        AGWN_HIP(hipEventRecord(e0, B.bs));
        AGWN_TRY(batch_launches(B, pos, x0f, variant, leaf, out, phase));
        AGWN_HIP(hipEventRecord(e1, B.bs));
        AGWN_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        AGWN_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms_out = double(ms);
        return ok;
    };
    int best = 0;
    double best_ms = 1e300;
    for (int v = 0; v < frac_batch_variants(); ++v)
    {
        AGWN_TRY(batch_prepare(B, pos, qoff, x0f, v, sample));
        if (v == 0)
            AGWN_TRY(batch_launches(B, pos, x0f, v, 0, out, kPhaseCount)); // the counts the fractional kernel reads
        double ms = 0.0;
        AGWN_TRY(timed(v, 0, kPhaseFrac, ms));
        if (v < int(sizeof(B.tune_ms) / sizeof(B.tune_ms[0])))
            B.tune_ms[v] = ms;
        if (ms < best_ms)
        {
            best_ms = ms;
            best = v;
        }
    }
    B.variant = best;
    AGWN_TRY(batch_prepare(B, pos, qoff, x0f, best, {}));
    int best_leaf = 0;
    best_ms = 1e300;
    for (int l = 0; l < kNumLeafSizes && B.ncitems > 0; ++l)
    {
        double ms = 0.0;
        AGWN_TRY(timed(best, l, kPhaseCount, ms));
        B.leaf_ms[l] = ms;
        if (ms < best_ms)
        {
            best_ms = ms;
            best_leaf = l;
        }
    }
    B.leaf = best_leaf;
    B.g_pos = nullptr; // the lists are rebuilt (and the graph captured) by the next evaluation
    return ok;
}

int batch_eval(agwn_batch* b, float const* pos, int64_t const* qoff, float const* x0, float* out, hipStream_t s)
{
    AGWN_CHECK(b != nullptr, "batch handle is null");
    AGWN_CHECK(qoff && x0, "null argument");
    AGWN_CHECK(qoff[0] == 0, "query offsets must start at 0");
    for (int64_t m = 0; m < b->k; ++m)
        AGWN_CHECK(qoff[m + 1] >= qoff[m], "query offsets must be non-decreasing (mesh %lld)", (long long)m);
    int64_t const n = qoff[b->k];
    AGWN_CHECK(n == 0 || (pos && out), "null query or output array");
    std::lock_guard<std::mutex> lk(b->mu);
    agwn_batch& B = *b;
    // order the batch's stream after the caller's work on `s`
    hipEvent_t ev = nullptr;
    AGWN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    struct evguard
    {
        hipEvent_t e;
        ~evguard() { (void)hipEventDestroy(e); }
    } eg{ev};
    AGWN_HIP(hipEventRecord(ev, s));
    AGWN_HIP(hipStreamWaitEvent(B.bs, ev, 0));
    if (n == 0)
        return ok;
    if (B.variant < 0)
        AGWN_TRY(batch_tune(B, pos, qoff, x0, out));
    std::vector<int64_t> q(qoff, qoff + B.k + 1);
    bool same = B.exec && B.g_pos == pos && B.g_out == out && B.g_x0[0] == x0[0] && B.g_x0[1] == x0[1] &&
                B.g_x0[2] == x0[2] && B.g_qoff == q && B.g_variant == B.variant && B.g_leaf == B.leaf;
    if (same)
    {
        // the graph's prepared BVH states must still cover the queries (their contents may have changed in place)
        std::vector<float> qmax;
        AGWN_TRY(batch_qmax(B, pos, qoff, qmax));
        for (int64_t m = 0; m < B.k && same; ++m)
        {
            mesh_t<float> const& M = *B.meshes[size_t(m)];
            if (qoff[m + 1] > qoff[m] && M.ntri > 0)
                same = B.states[size_t(m)] && double(qmax[size_t(m)]) + M.vmax <= B.states[size_t(m)]->dsum;
        }
    }
    if (!same)
    {
        if (B.exec)
        {
            AGWN_HIP(hipGraphExecDestroy(B.exec));
            B.exec = nullptr;
        }
        AGWN_TRY(batch_prepare(B, pos, qoff, x0, B.variant, {}));
        hipGraph_t graph = nullptr;
        AGWN_HIP(hipStreamBeginCapture(B.bs, hipStreamCaptureModeThreadLocal));
        int const st = batch_launches(B, pos, x0, B.variant, B.leaf, out);
        hipError_t const ce = hipStreamEndCapture(B.bs, &graph);
        AGWN_TRY(st);
        AGWN_HIP(ce);
        hipError_t const ie = hipGraphInstantiate(&B.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        AGWN_HIP(ie);
        B.g_pos = pos;
        B.g_out = out;
        B.g_x0[0] = x0[0];
        B.g_x0[1] = x0[1];
        B.g_x0[2] = x0[2];
        B.g_qoff = q;
        B.g_variant = B.variant;
        B.g_leaf = B.leaf;
    }
    AGWN_HIP(hipGraphLaunch(B.exec, B.bs));
    AGWN_HIP(hipStreamSynchronize(B.bs));
    return ok;
}
} // namespace
} // namespace agwn

extern "C" {
int agwn_batch_create(int64_t k, float const* verts, int64_t const* vert_off, int32_t const* idx,
                      int64_t const* idx_off, float const* seg7, int64_t const* seg_off, agwn_batch** out)
{
    return guarded2([&] { return batch_create(k, verts, vert_off, idx, idx_off, seg7, seg_off, out); });
}
int agwn_batch_destroy(agwn_batch* b)
{
    clear_error();
    delete b;
    return ok;
}
int agwn_batch_tune(agwn_batch* b, float const* pos, int64_t const* q_off, float const* x0, float* out,
                    agwn_stream s)
{
    return guarded2([&]() -> int {
        AGWN_CHECK(b != nullptr, "batch handle is null");
        AGWN_CHECK(q_off && x0, "null argument");
        std::lock_guard<std::mutex> lk(b->mu);
        AGWN_HIP(hipStreamSynchronize(hipStream_t(s)));
        return batch_tune(*b, pos, q_off, x0, out);
    });
}
int agwn_batch_eval_dev(agwn_batch* b, float const* pos, int64_t const* q_off, float const* x0, float* out,
                        agwn_stream s)
{
    return guarded2([&] { return batch_eval(b, pos, q_off, x0, out, hipStream_t(s)); });
}
int agwn_batch_info(agwn_batch const* b, int64_t* k, int64_t* ntri, int64_t* nunit, int* frac_shape, int* leaf,
                    double* frac_ms, double* leaf_ms)
{
    return guarded2([&]() -> int {
        AGWN_CHECK(b != nullptr, "batch handle is null");
        int64_t t = 0, u = 0;
        for (int64_t m = 0; m < b->k; ++m)
        {
            t += b->meshes[size_t(m)]->ntri;
            u += b->nunit[size_t(m)];
        }
        if (k)
            *k = b->k;
        if (ntri)
            *ntri = t;
        if (nunit)
            *nunit = u;
        if (frac_shape)
            *frac_shape = b->variant;
        if (leaf)
            *leaf = b->leaf;
        if (frac_ms)
            for (int v = 0; v < frac_batch_variants(); ++v)
                frac_ms[v] = b->tune_ms[v];
        if (leaf_ms)
            for (int l = 0; l < kNumLeafSizes; ++l)
                leaf_ms[l] = b->leaf_ms[l];
        return ok;
    });
}
int agwn_batch_set_tuning(agwn_batch* b, int frac_shape, int leaf)
{
    return guarded2([&]() -> int {
        AGWN_CHECK(b != nullptr, "batch handle is null");
        AGWN_CHECK(frac_shape >= 0 && frac_shape < frac_batch_variants(), "fractional shape %d out of range [0, %d)",
                   frac_shape, frac_batch_variants());
        AGWN_CHECK(leaf >= 0 && leaf < kNumLeafSizes, "leaf-bucket size index %d out of range [0, %d)", leaf,
                   kNumLeafSizes);
        std::lock_guard<std::mutex> lk(b->mu);
        b->variant = frac_shape;
        b->leaf = leaf;
        return ok;
    });
}
int agwn_batch_frac_shape(int v, int* frac_q, int* frac_block, int* frac_stage)
{
    return guarded2([&]() -> int {
        AGWN_CHECK(v >= 0 && v < frac_batch_variants(), "fractional shape %d out of range [0, %d)", v,
                   frac_batch_variants());
        int q = 0, blk = 0, st = 0;
        frac_batch_shape(v, q, blk, st);
        if (frac_q)
            *frac_q = q;
        if (frac_block)
            *frac_block = blk;
        if (frac_stage)
            *frac_stage = st;
        return ok;
    });
}
int agwn_batch_leaf_size(int l, int* leaf)
{
    return guarded2([&]() -> int {
        AGWN_CHECK(l >= 0 && l < kNumLeafSizes, "leaf-bucket size index %d out of range [0, %d)", l, kNumLeafSizes);
        if (leaf)
            *leaf = kLeafSizes[l];
        return ok;
    });
}
} // extern "C"
