// Robust GWN on gfx950: RobustMeshGwn<T, Precision> (gwn_mesh_robust.hh:170-456) as a prepared device object.
//
// Precision policies (robust_predicates.hh:186-215), template parameter C of every kernel:
//   int32_t  int64_grid   20-bit quantized coordinates, 64-bit determinants (exact)
//   int64_t  int128_grid  50-bit quantized coordinates, 128-bit determinants (exact; __int128 lowered to 64-bit ops)
//   double   exact_float  the doubles themselves; determinant signs exact through a floating-point filter and, when
//                         it cannot decide, an exact expansion (two-product via FMA + expansion growth, Shewchuk 1997)
// Every sign the count and the fractional term depend on is computed exactly, so the integer count is bit-exact with
// the reference's and the fractional term differs from it only by the last bits of FP64 sqrt/atan2.
//
// Integer term: the -x ray from the quantized query, culled by an LBVH whose leaf boxes are the exact bounding boxes
// of the predicate's own triangles (in predicate coordinates, so the (y,z) test is exact), grown by a few units of
// rounding of the FP64 front/back test hit_x (a barycentric mix of the vertex depths, which may round just outside
// them); a box is a candidate when lo.y <= q.y <= hi.y, lo.z <= q.z <= hi.z and lo.x <= q.x. This replaces Embree's
// user-geometry traversal (gwn_mesh_robust.hh:91-145, 380-407): both only cull what the exact predicate rejects.
// Fractional term: one lane per query walks the swapped, quantized boundary in order (gwn_mesh_robust.hh:264-325).

#include "../../include/agwn.h"
#include "agwn_lbvh.hpp"
#include "agwn_math.hpp"

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <memory>
#include <type_traits>
#include <unordered_map>
#include <vector>

namespace agwn
{
namespace
{
constexpr int kB = 256;
constexpr int kTB = 128;
constexpr double kPi = 3.141592653589793238462643383279502884;

using i128 = __int128;

template <class C>
struct wide_t_of;
template <>
struct wide_t_of<int32_t>
{
    using type = int64_t;
};
template <>
struct wide_t_of<int64_t>
{
    using type = i128;
};
template <>
struct wide_t_of<double>
{
    using type = double;
};
template <class C>
using wide = typename wide_t_of<C>::type;

template <class C>
struct v3
{
    C x, y, z;
};

// correctly rounded (nearest-even) conversion of a signed 128-bit integer to double
__host__ __device__ inline double i128_to_double(i128 v)
{
    bool const neg = v < 0;
    unsigned __int128 m = neg ? (unsigned __int128)0 - (unsigned __int128)v : (unsigned __int128)v;
    uint64_t const hi = uint64_t(m >> 64);
    double r;
    if (hi == 0)
    {
        r = double(uint64_t(m));
    }
    else
    {
#if defined(__HIP_DEVICE_COMPILE__)
        int const lz = __clzll((long long)hi);
#else
        int const lz = __builtin_clzll(hi);
#endif
        int const shift = 64 - lz; // bits above the low 64 that must be dropped
        uint64_t t = uint64_t(m >> shift);
        unsigned __int128 const rest = m & (((unsigned __int128)1 << shift) - 1);
        t |= rest != 0 ? 1u : 0u; // sticky bit: keeps nearest-even rounding of the 64-bit value exact
        r = ldexp(double(t), shift);
    }
    return neg ? -r : r;
}

template <class W>
__host__ __device__ inline double to_d(W w)
{
    if constexpr (std::is_same_v<W, i128>)
        return i128_to_double(w);
    else
        return double(w);
}

template <class C>
__host__ __device__ inline wide<C> wmul(C a, C b)
{
    return wide<C>(a) * wide<C>(b);
}

// --- exact sign of a 6-product sum of doubles (expansion arithmetic) ------------------------------------------------
__device__ inline void two_sum(double a, double b, double& s, double& e)
{
#pragma clang fp contract(off)
    s = a + b;
    double const bv = s - a;
    double const av = s - bv;
    e = (a - av) + (b - bv);
}

// sign of sum_i a[i] * b[i], exactly: each product is split into p + err (err = fma(a, b, -p), exact), and all parts
// are accumulated into a nonoverlapping expansion whose most significant nonzero component carries the sign.
__device__ inline int exact_sign6(double const (&a)[6], double const (&b)[6])
{
#pragma clang fp contract(off)
    double e[12];
    int n = 0;
    for (int i = 0; i < 6; ++i)
    {
        double const p = a[i] * b[i];
        double const parts[2] = {fma(a[i], b[i], -p), p};
        for (double v : parts)
        {
            for (int k = 0; k < n; ++k)
            {
                double s, err;
                two_sum(v, e[k], s, err);
                e[k] = err;
                v = s;
            }
            e[n++] = v;
        }
    }
    for (int k = n - 1; k >= 0; --k)
        if (e[k] != 0.0)
            return e[k] > 0.0 ? 1 : -1;
    return 0;
}

// robust_predicates.hh:272-308 -- edge_det: 2x2 determinant of (q - q0, q1 - q0) in (y, z)
template <class C>
__device__ inline wide<C> edge_det(v3<C> q, v3<C> q0, v3<C> q1)
{
#pragma clang fp contract(off)
    if constexpr (std::is_same_v<C, double>)
    {
        double const p1 = (q.y - q0.y) * (q0.z - q1.z);
        double const p2 = (q.z - q0.z) * (q1.y - q0.y);
        double const det = p1 + p2;
        constexpr double eps = DBL_EPSILON / 2;
        constexpr double err_bound = (3.0 + 16.0 * eps) * eps;
        double const mag = fabs(p1) + fabs(p2);
        if (fabs(det) > err_bound * mag)
            return det;
        if (mag == 0.0)
            return 0.0;
        double const a[6] = {q.y, -q.y, q0.y, q.z, -q.z, -q0.z};
        double const b[6] = {q0.z, q1.z, q1.z, q1.y, q0.y, q1.y};
        int const sgn = exact_sign6(a, b);
        if (sgn == 0)
            return 0.0;
        return copysign(fmax(fabs(det), DBL_MIN), double(sgn));
    }
    else
    {
        return wmul<C>(C(q.y - q0.y), C(q0.z - q1.z)) + wmul<C>(C(q.z - q0.z), C(q1.y - q0.y));
    }
}

// robust_predicates.hh:314-331 -- edge_sign with the (eps2 on y) >> (eps3 on z) perturbation
template <class C>
__device__ inline int edge_sign(v3<C> q, v3<C> q0, v3<C> q1)
{
    wide<C> const t = edge_det(q, q0, q1);
    if (t != 0)
        return t > 0 ? 1 : -1;
    if (q0.z != q1.z)
        return q0.z > q1.z ? 1 : -1;
    if (q1.y != q0.y)
        return q1.y > q0.y ? 1 : -1;
    return 0;
}

// robust_predicates.hh:339-356 -- atan2 numerator with exact sign (== -edge_det)
template <class C>
__device__ inline wide<C> atan2_num(v3<C> q, v3<C> q0, v3<C> q1)
{
#pragma clang fp contract(off)
    v3<C> const v0{C(q0.x - q.x), C(q0.y - q.y), C(q0.z - q.z)};
    v3<C> const v1{C(q1.x - q.x), C(q1.y - q.y), C(q1.z - q.z)};
    if constexpr (std::is_same_v<C, double>)
    {
        double const t = edge_det(q, q0, q1);
        if (t == 0.0)
            return 0.0;
        double const num = v0.z * v1.y - v0.y * v1.z;
        return copysign(fmax(fabs(num), DBL_MIN), -t);
    }
    else
    {
        return wmul<C>(v0.z, v1.y) - wmul<C>(v0.y, v1.z);
    }
}

// robust_predicates.hh:368-375 -- dot_yz
template <class C>
__device__ inline double dot_yz(v3<C> a, v3<C> b)
{
#pragma clang fp contract(off)
    if constexpr (std::is_same_v<C, double>)
        return a.y * b.y + a.z * b.z;
    else
        return to_d(wmul<C>(a.y, b.y) + wmul<C>(a.z, b.z));
}

// --- quantization (gwn_mesh_robust.hh:334-351) ----------------------------------------------------------------------
struct quant
{
    double cx, cy, cz, scale;
    int grid_bits; // 0: exact_float
};

template <class C>
__host__ __device__ inline C qround(double v, double max_q)
{
    double const c = v < -max_q ? -max_q : (max_q < v ? max_q : v);
    return C(llround(c));
}

template <class C>
__host__ __device__ inline v3<C> quantize(quant const& Q, double x, double y, double z)
{
    if constexpr (std::is_same_v<C, double>)
    {
        return {x, y, z};
    }
    else
    {
#pragma clang fp contract(off)
        double const max_q = double(int64_t(1) << (Q.grid_bits + 10));
        double const qx = (x - Q.cx) * Q.scale;
        double const qy = (y - Q.cy) * Q.scale;
        double const qz = (z - Q.cz) * Q.scale;
        return {qround<C>(qx, max_q), qround<C>(qy, max_q), qround<C>(qz, max_q)};
    }
}

template <class C>
struct alignas(sizeof(C) == 4 ? 32 : 64) rseg
{
    v3<C> p0, p1; // stored swapped: p0 = quantize(pos1), p1 = quantize(pos0) (gwn_mesh_robust.hh:198-206)
    double w;
};

// --- device construction (gwn_mesh_robust.hh:177-209, build_quantization :354-378) -----------------------------------
// Exact bounding box: doubles mapped to order-preserving uint64 keys, reduced per lane over a strided range, then one
// atomic min / max per lane. box[0..2] = min keys, box[3..5] = max keys.
__device__ __forceinline__ unsigned long long dkey(double v)
{
    unsigned long long const b = static_cast<unsigned long long>(__double_as_longlong(v));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
inline double dkey_value(unsigned long long k)
{
    unsigned long long const b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    std::memcpy(&v, &b, sizeof(v));
    return v;
}
template <class T>
__global__ void __launch_bounds__(kB) k_rbox(T const* __restrict__ v, int64_t nv, unsigned long long* __restrict__ box)
{
    unsigned long long mn[3] = {~0ull, ~0ull, ~0ull}, mx[3] = {0ull, 0ull, 0ull};
    for (int64_t i = int64_t(blockIdx.x) * kB + threadIdx.x; i < nv; i += int64_t(gridDim.x) * kB)
        for (int k = 0; k < 3; ++k)
        {
            unsigned long long const key = dkey(double(v[3 * i + k]));
            mn[k] = key < mn[k] ? key : mn[k];
            mx[k] = key > mx[k] ? key : mx[k];
        }
    for (int k = 0; k < 3; ++k)
    {
        atomicMin(box + k, mn[k]);
        atomicMax(box + 3 + k, mx[k]);
    }
}
// quantized triangle corners (input order) and swapped boundary segments
template <class C, class T>
__global__ void __launch_bounds__(kB) k_rquant_tri(T const* __restrict__ v, int32_t const* __restrict__ idx,
                                                   int64_t n3, quant Q, v3<C>* __restrict__ out)
{
    int64_t const t = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (t >= n3)
        return;
    T const* p = v + 3 * int64_t(idx[t]);
    out[t] = quantize<C>(Q, double(p[0]), double(p[1]), double(p[2]));
}
template <class C, class T>
__global__ void __launch_bounds__(kB) k_rquant_seg(T const* __restrict__ seg7, int64_t ns, quant Q,
                                                   rseg<C>* __restrict__ out)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= ns)
        return;
    T const* g = seg7 + 7 * k;
    rseg<C> r;
    r.p0 = quantize<C>(Q, double(g[3]), double(g[4]), double(g[5]));
    r.p1 = quantize<C>(Q, double(g[0]), double(g[1]), double(g[2]));
    r.w = double(g[6]);
    out[k] = r;
}

// --- integer term ---------------------------------------------------------------------------------------------------
template <class C>
__device__ inline int leaf_count(v3<C> qi, v3<C> const* tri)
{
#pragma clang fp contract(off)
    v3<C> const a = tri[0], b = tri[1], c = tri[2];
    // robust_predicates.hh:390-402 (project_into_tri)
    int const e01 = edge_sign(qi, a, b);
    int const e12 = edge_sign(qi, b, c);
    int const e20 = edge_sign(qi, c, a);
    if (e01 == 0 || e01 != e12 || e01 != e20)
        return 0;
    double const w0 = to_d(edge_det(qi, b, c));
    double const w1 = to_d(edge_det(qi, c, a));
    double const w2 = to_d(edge_det(qi, a, b));
    // gwn_mesh_robust.hh:141-144
    double const hit_x = (w0 * double(a.x) + w1 * double(b.x) + w2 * double(c.x)) / (w0 + w1 + w2);
    return double(qi.x) >= hit_x ? -e01 : 0;
}

__device__ __forceinline__ bool rbox(double const* lo, double const* hi, double qx, double qy, double qz)
{
    return lo[1] <= qy && qy <= hi[1] && lo[2] <= qz && qz <= hi[2] && lo[0] <= qx;
}

template <class C, class T>
__global__ void __launch_bounds__(kTB)
    k_rcount(T const* __restrict__ pos, int64_t n, quant Q, bvh_node<double> const* __restrict__ nodes,
             v3<C> const* __restrict__ tri, int64_t ntri, int32_t* __restrict__ out)
{
    __shared__ int32_t stack[kMaxStack][kTB];
    int64_t const i = int64_t(blockIdx.x) * kTB + threadIdx.x;
    if (i >= n)
        return;
    v3<C> const qi = quantize<C>(Q, double(pos[3 * i]), double(pos[3 * i + 1]), double(pos[3 * i + 2]));
    double const qx = double(qi.x), qy = double(qi.y), qz = double(qi.z);
    int sum = 0;
    if (ntri == 1)
    {
        sum = leaf_count(qi, tri);
    }
    else if (ntri > 1)
    {
        int sp = 0;
        int32_t node = 0;
        while (true)
        {
            bvh_node<double> const nd = nodes[node];
            bool const h0 = rbox(nd.lo0, nd.hi0, qx, qy, qz);
            bool const h1 = rbox(nd.lo1, nd.hi1, qx, qy, qz);
            int32_t next = -1;
            bool have = false;
            if (h0)
            {
                if (nd.c0 < 0)
                    sum += leaf_count(qi, tri + 3 * int64_t(~nd.c0));
                else
                {
                    next = nd.c0;
                    have = true;
                }
            }
            if (h1)
            {
                if (nd.c1 < 0)
                    sum += leaf_count(qi, tri + 3 * int64_t(~nd.c1));
                else if (have)
                {
                    // both children internal: the smaller subtree first, the larger one on the stack (<= log2 n)
                    int32_t const push = nd.big ? nd.c1 : nd.c0;
                    next = nd.big ? nd.c0 : nd.c1;
                    if (sp < kMaxStack)
                        stack[sp++][threadIdx.x] = push;
                }
                else
                {
                    next = nd.c1;
                    have = true;
                }
            }
            if (have)
                node = next;
            else if (sp == 0)
                break;
            else
                node = stack[--sp][threadIdx.x];
        }
    }
    out[i] = sum;
}

// --- fractional term (gwn_mesh_robust.hh:264-325) -------------------------------------------------------------------
// The boundary as a chain stream. Stored segments (p0, p1) are linked where one's p1 is the next one's p0 (bit for bit),
// so consecutive stream points are the endpoints of a segment and each point's per-query quantities (v = P - q,
// l = |v|, l + v.x) are computed once for the two segments that share it. seg = 1: the segment (previous point, this
// point) of weight w ends here; seg = 0: a walk starts here. The reference sums the segments' terms in input order;
// the stream sums the same terms in walk order.
template <class C>
struct rpt
{
    v3<C> p;
    int32_t seg;
    double w;
};

// one stream point's per-query quantities
template <class C>
struct rpoint
{
    v3<C> p, v;          // the point and P - q, in predicate coordinates
    double dx, dy, dz;   // v in FP64
    double lpx;          // l + v.x without cancellation (l_plus_x)
};

template <class C>
__device__ inline double l_plus_x(v3<C> v, double l)
{
#pragma clang fp contract(off)
    if (v.x >= 0)
        return l + double(v.x);
    return dot_yz(v, v) / (l - double(v.x));
}

// F = false: the reference's operations (gwn_mesh_robust.hh:264-290). F = true (int64_grid only): the same values in
// FP64. Mesh coordinates are quantized to |c| < 2^20 and a query is clamped to 2^30; for a query with |q| <= 2^25
// every coordinate difference is below 2^25 + 2^20 < 2^25.1 in magnitude, so every product of two differences is
// below 2^50.2, every sum of two products below 2^51.2 and a squared length below 2^51.8: integers that double
// represents exactly. The 64-bit integer numerator, squared lengths and dot products of the reference
// (robust_predicates.hh:339-375) then have the same values in FP64, and only sqrt and the division of l_plus_x round
// (within ~1 ulp, where the reference's are correctly rounded).
template <bool F, class C>
__device__ inline rpoint<C> rpoint_at(v3<C> qi, v3<C> P)
{
#pragma clang fp contract(off)
    rpoint<C> r;
    r.p = P;
    r.v = v3<C>{C(P.x - qi.x), C(P.y - qi.y), C(P.z - qi.z)};
    r.dx = double(r.v.x);
    r.dy = double(r.v.y);
    r.dz = double(r.v.z);
    if constexpr (F)
    {
        static_assert(std::is_same_v<C, int32_t>, "the FP64 form is int64_grid's");
        double const yz = __builtin_fma(r.dy, r.dy, r.dz * r.dz);
        double const l = sqrt_f64_normal(__builtin_fma(r.dx, r.dx, yz));
        r.lpx = r.dx >= 0.0 ? l + r.dx : div_f64_normal(yz, l - r.dx);
    }
    else
    {
        double const l = sqrt(r.dx * r.dx + r.dy * r.dy + r.dz * r.dz);
        r.lpx = l_plus_x(r.v, l);
    }
    return r;
}

// The term of the segment (a, b): (num, den) of its atan2 (the generic case, num != 0: returns false), or the term
// itself in `special` (num == 0: the ray through an endpoint, grazing: returns true). The reference's den cancels for
// a query next to the segment, so its rounding is kept as is.
template <bool F, class C>
__device__ inline bool rterm(v3<C> qi, rpoint<C> const& a, rpoint<C> const& b, double& num, double& den,
                             double& special)
{
#pragma clang fp contract(off)
    bool nonzero;
    if constexpr (F)
    {
        den = a.lpx * b.lpx + __builtin_fma(a.dy, b.dy, a.dz * b.dz);
        num = __builtin_fma(a.dz, b.dy, -(a.dy * b.dz));
        nonzero = num != 0.0;
    }
    else
    {
        wide<C> const wn = atan2_num(qi, a.p, b.p);
        den = a.lpx * b.lpx + dot_yz(a.v, b.v);
        nonzero = wn != 0;
        if (nonzero)
            num = to_d(wn);
    }
    if (nonzero)
        return false;
    v3<C> const v0 = a.v, v1 = b.v;
    if (v0.y == 0 && v0.z == 0 && v0.x < 0 && (v1.y != 0 || v1.z != 0))
        special = v1.z != 0 ? atan2_f64(b.dz, -b.dy) : (v1.y > 0 ? -kPi : 0.0);
    else if (v1.y == 0 && v1.z == 0 && v1.x < 0 && (v0.y != 0 || v0.z != 0))
        special = v0.z != 0 ? atan2_f64(-a.dz, -a.dy) : (v0.y > 0 ? kPi : 0.0);
    else
        special = den < 0.0 ? double(-edge_sign(qi, a.p, b.p)) * kPi : 0.0;
    return true;
}

// One lane per query walks the chain stream, staged through LDS a tile at a time (broadcast reads). Consecutive
// segments pair up: two generic terms of equal weight whose den are positive and whose magnitudes lie in
// [2^-400, 2^400] share one atan2 (the argument of the product of their complex numbers is the exact sum of two angles
// in (-pi/2, pi/2)); any other term takes its own. The arctangent is atan2_f64 (agwn_math.hpp), within 2 ulp of the
// library's. int64_grid takes the FP64 form when every query of the workgroup lies within 2^25 (the choice is
// uniform over the workgroup, whose waves share the LDS tiles).
template <class C, class T>
__global__ void __launch_bounds__(kB)
    k_rfrac(T const* __restrict__ pos, int64_t n, quant Q, rpt<C> const* __restrict__ pts, int64_t npts,
            double f_norm, int32_t const* __restrict__ counts, T* __restrict__ out_frac, T* __restrict__ out_eval)
{
#pragma clang fp contract(off)
    __shared__ rpt<C> tile[kB];
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    double x = 0.0, y = 0.0, z = 0.0;
    if (i < n)
    {
        x = double(pos[3 * i]);
        y = double(pos[3 * i + 1]);
        z = double(pos[3 * i + 2]);
    }
    v3<C> const qi = quantize<C>(Q, x, y, z);
    constexpr double kLo = 0x1p-400, kHi = 0x1p400;
    auto single = [](bool sp, double nu, double de, double spv) { return sp ? spv : atan2_f64(nu, de); };
    double half_area = 0.0;
    auto walk = [&](auto f_tag) {
        constexpr bool F = decltype(f_tag)::value;
        rpoint<C> a{}; // the previous stream point
        double pn = 0.0, pd = 0.0, ps = 0.0, pw = 0.0; // the pending term
        bool psp = false, pending = false;               // pending is the same in every lane
        for (int64_t t0 = 0; t0 < npts; t0 += kB)
        {
            __syncthreads(); // the previous tile has been consumed by every wave
            if (t0 + threadIdx.x < npts)
                tile[threadIdx.x] = pts[t0 + threadIdx.x];
            __syncthreads();
            int const cnt = int(npts - t0 < kB ? npts - t0 : kB);
            for (int j = 0; j < cnt; ++j)
            {
                rpt<C> const g = tile[j];
                rpoint<C> const b = rpoint_at<F>(qi, g.p);
                if (__builtin_amdgcn_readfirstlane(g.seg) != 0)
                {
                    double nu = 0.0, de = 0.0, sv = 0.0;
                    bool const sp = rterm<F>(qi, a, b, nu, de, sv);
                    double const w = g.w;
                    if (pending)
                    {
                        // F: numerators are nonzero integers below 2^51.2 and den below 2^56, always in range
                        bool in_range = true;
                        if constexpr (!F)
                        {
                            double const m0 = __builtin_fmax(__builtin_fabs(pn), __builtin_fabs(pd));
                            double const m1 = __builtin_fmax(__builtin_fabs(nu), __builtin_fabs(de));
                            in_range = m0 >= kLo && m0 <= kHi && m1 >= kLo && m1 <= kHi;
                        }
                        if (!psp && !sp && pw == w && pd > 0.0 && de > 0.0 && in_range)
                            half_area += w * atan2_f64(__builtin_fma(pn, de, nu * pd), __builtin_fma(pd, de, -(pn * nu)));
                        else
                        {
                            half_area += pw * single(psp, pn, pd, ps);
                            half_area += w * single(sp, nu, de, sv);
                        }
                    }
                    else
                    {
                        pn = nu;
                        pd = de;
                        ps = sv;
                        psp = sp;
                        pw = w;
                    }
                    pending = !pending;
                }
                a = b;
            }
        }
        if (pending)
            half_area += pw * single(psp, pn, pd, ps);
    };
    if constexpr (std::is_same_v<C, int32_t>)
    {
        int32_t const qm = max(max(abs(qi.x), abs(qi.y)), abs(qi.z));
        if (__syncthreads_and(i >= n || qm <= (1 << 25)))
            walk(std::true_type{});
        else
            walk(std::false_type{});
    }
    else
    {
        walk(std::false_type{});
    }
    if (i < n)
    {
        T const f = T(half_area * f_norm);
        if (out_frac)
            out_frac[i] = f;
        if (out_eval)
            out_eval[i] = f + T(counts[i]);
    }
}

template <class C>
__global__ void __launch_bounds__(kB)
    k_rleaf_boxes(v3<C> const* __restrict__ tri, int64_t ntri, double grid_pad, double* __restrict__ lo,
                  double* __restrict__ hi, float* __restrict__ cent)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= ntri)
        return;
    v3<C> const a = tri[3 * k], b = tri[3 * k + 1], c = tri[3 * k + 2];
    double const xs[3] = {double(a.x), double(b.x), double(c.x)};
    double const ys[3] = {double(a.y), double(b.y), double(c.y)};
    double const zs[3] = {double(a.z), double(b.z), double(c.z)};
    double const* cs[3] = {xs, ys, zs};
    double mx_abs = 0.0;
    for (int ax = 0; ax < 3; ++ax)
        for (int j = 0; j < 3; ++j)
            mx_abs = fmax(mx_abs, fabs(cs[ax][j]));
    // rounding of the FP64 front/back test (a barycentric mix of the depths): a few units in the last place
    double const pad = 16.0 * DBL_EPSILON * mx_abs + grid_pad + 4.0 * DBL_MIN;
    for (int ax = 0; ax < 3; ++ax)
    {
        double const mn = fmin(fmin(cs[ax][0], cs[ax][1]), cs[ax][2]);
        double const mx = fmax(fmax(cs[ax][0], cs[ax][1]), cs[ax][2]);
        if (lo)
        {
            lo[3 * k + ax] = mn - pad;
            hi[3 * k + ax] = mx + pad;
        }
        if (cent)
            cent[3 * k + ax] = float(0.5 * (mn + mx));
    }
}

template <class C>
__global__ void __launch_bounds__(kB) k_rgather(v3<C> const* __restrict__ src, int32_t const* __restrict__ order,
                                                int64_t ntri, v3<C>* __restrict__ dst)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= ntri)
        return;
    int64_t const t = order[k];
    for (int j = 0; j < 3; ++j)
        dst[3 * k + j] = src[3 * t + j];
}
// Links the quantized segments into walks for the chain stream: a walk starts at the first unused segment (input
// order) and continues with an unused segment whose p0 equals its last p1 bit for bit, while there is one. Every
// segment appears once; a closed loop becomes one walk of n + 1 points.
template <class C>
std::vector<rpt<C>> chain_stream(std::vector<rseg<C>> const& sg)
{
    struct key
    {
        uint64_t k[3];
        bool operator==(key const& o) const { return k[0] == o.k[0] && k[1] == o.k[1] && k[2] == o.k[2]; }
    };
    struct key_hash
    {
        size_t operator()(key const& x) const
        {
            uint64_t h = 0x9e3779b97f4a7c15ull;
            for (uint64_t v : x.k)
                h = (h ^ v) * 0xff51afd7ed558ccdull, h ^= h >> 32;
            return size_t(h);
        }
    };
    auto key_of = [](v3<C> const& p) {
        key r{};
        C const c[3] = {p.x, p.y, p.z};
        for (int j = 0; j < 3; ++j)
            std::memcpy(&r.k[j], &c[j], sizeof(C));
        return r;
    };
    int64_t const ns = int64_t(sg.size());
    std::unordered_map<key, std::vector<int64_t>, key_hash> from; // segments by their p0, each list descending
    from.reserve(size_t(ns));
    for (int64_t k = ns - 1; k >= 0; --k)
        from[key_of(sg[size_t(k)].p0)].push_back(k);
    std::vector<char> used(size_t(ns), 0);
    std::vector<rpt<C>> out;
    out.reserve(size_t(ns) + 16);
    for (int64_t k = 0; k < ns; ++k)
    {
        if (used[size_t(k)])
            continue;
        out.push_back(rpt<C>{sg[size_t(k)].p0, 0, 0.0});
        int64_t cur = k;
        while (cur >= 0)
        {
            used[size_t(cur)] = 1;
            rseg<C> const& g = sg[size_t(cur)];
            out.push_back(rpt<C>{g.p1, 1, g.w});
            cur = -1;
            auto it = from.find(key_of(g.p1));
            if (it == from.end())
                break;
            std::vector<int64_t>& cand = it->second;
            while (!cand.empty() && used[size_t(cand.back())])
                cand.pop_back();
            if (!cand.empty())
            {
                cur = cand.back();
                cand.pop_back();
            }
        }
    }
    return out;
}
} // namespace

// ---------------------------------------------------------------------------------------------------------------------
// prepared object
// ---------------------------------------------------------------------------------------------------------------------
struct robust_impl
{
    virtual ~robust_impl() = default;
    int scalar_bytes = 0;
    int precision = 0;
    int64_t ntri = 0;
    int64_t nseg = 0;
    quant Q{0, 0, 0, 1.0, 0};
    virtual int count(void const* pos, int64_t n, int32_t* out, hipStream_t s) const = 0;
    virtual int frac(void const* pos, int64_t n, int32_t const* counts, void* out_frac, void* out_eval,
                     hipStream_t s) const = 0;
};

namespace
{
template <class C, class T>
struct robust_t final : robust_impl
{
    v3<C>* tri = nullptr;     // predicate triangles in LBVH leaf order
    rpt<C>* stream = nullptr; // the swapped quantized boundary as a chain stream
    int64_t nstream = 0;
    bvh_node<double>* nodes = nullptr;
    lbvh_topology topo;

    ~robust_t() override
    {
        for (void* p : {(void*)tri, (void*)stream, (void*)nodes})
            if (p)
                (void)hipFree(p);
        lbvh_release(topo);
    }

    // gwn_mesh_robust.hh:177-209 + build_quantization :354-378, on the device: the inputs are uploaded once, the bounding
    // box is reduced by k_rbox, the quantization parameters follow from it on the host with the reference's
    // arithmetic, and triangles and segments are quantized by kernels (the queries' quantize(), contraction off)
    int create(T const* hv, int64_t nv, int32_t const* hidx, int64_t nidx, T const* hseg, int64_t ns, int bits)
    {
#pragma clang fp contract(off)
        ntri = nidx / 3;
        nseg = ns;
        Q.grid_bits = bits;
        hipStream_t s = nullptr;
        AGWN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        struct sguard
        {
            hipStream_t s;
            ~sguard() { (void)hipStreamDestroy(s); }
        } sg{s};
        dbuf dv, di, ds;
        AGWN_TRY(dv.alloc(size_t(std::max<int64_t>(nv, 1)) * 3 * sizeof(T), s));
        AGWN_TRY(di.alloc(size_t(std::max<int64_t>(nidx, 1)) * sizeof(int32_t), s));
        AGWN_TRY(ds.alloc(size_t(std::max<int64_t>(ns, 1)) * 7 * sizeof(T), s));
        if (nv > 0)
            AGWN_HIP(hipMemcpyAsync(dv.p, hv, size_t(nv) * 3 * sizeof(T), hipMemcpyHostToDevice, s));
        if (nidx > 0)
            AGWN_HIP(hipMemcpyAsync(di.p, hidx, size_t(nidx) * sizeof(int32_t), hipMemcpyHostToDevice, s));
        if (ns > 0)
            AGWN_HIP(hipMemcpyAsync(ds.p, hseg, size_t(ns) * 7 * sizeof(T), hipMemcpyHostToDevice, s));
        if (bits > 0 && nv > 0)
        {
            dbuf db;
            AGWN_TRY(db.alloc(6 * sizeof(unsigned long long), s));
            unsigned long long const init[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
            AGWN_HIP(hipMemcpyAsync(db.p, init, sizeof(init), hipMemcpyHostToDevice, s));
            unsigned const gb = unsigned(std::min<int64_t>((nv + kB - 1) / kB, 1024));
            k_rbox<T><<<dim3(gb), dim3(kB), 0, s>>>(dv.as<T>(), nv, db.as<unsigned long long>());
            AGWN_HIP(hipGetLastError());
            unsigned long long hb[6];
            AGWN_HIP(hipMemcpyAsync(hb, db.p, sizeof(hb), hipMemcpyDeviceToHost, s));
            AGWN_HIP(hipStreamSynchronize(s));
            double const mn[3] = {dkey_value(hb[0]), dkey_value(hb[1]), dkey_value(hb[2])};
            double const mx[3] = {dkey_value(hb[3]), dkey_value(hb[4]), dkey_value(hb[5])};
            double const ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
            Q.cx = mn[0] + ex * 0.5;
            Q.cy = mn[1] + ey * 0.5;
            Q.cz = mn[2] + ez * 0.5;
            double const half = 0.5 * std::max({ex, ey, ez});
            double const max_coord = double((int64_t(1) << bits) - 1);
            Q.scale = half > 0.0 ? max_coord / half : 1.0;
        }
        if (ns > 0)
        {
            dbuf dq;
            AGWN_TRY(dq.alloc(size_t(ns) * sizeof(rseg<C>), s));
            k_rquant_seg<C, T><<<dim3(unsigned((ns + kB - 1) / kB)), dim3(kB), 0, s>>>(ds.as<T>(), ns, Q,
                                                                                      dq.as<rseg<C>>());
            AGWN_HIP(hipGetLastError());
            std::vector<rseg<C>> hq(static_cast<size_t>(ns));
            AGWN_HIP(hipMemcpyAsync(hq.data(), dq.p, size_t(ns) * sizeof(rseg<C>), hipMemcpyDeviceToHost, s));
            AGWN_HIP(hipStreamSynchronize(s));
            std::vector<rpt<C>> const hs = chain_stream(hq);
            nstream = int64_t(hs.size());
            AGWN_HIP(hipMalloc(&stream, hs.size() * sizeof(rpt<C>)));
            AGWN_HIP(hipMemcpyAsync(stream, hs.data(), hs.size() * sizeof(rpt<C>), hipMemcpyHostToDevice, s));
            AGWN_HIP(hipStreamSynchronize(s));
        }
        if (ntri > 0)
        {
            dbuf dt, dc;
            AGWN_TRY(dt.alloc(size_t(3 * ntri) * sizeof(v3<C>), s));
            AGWN_TRY(dc.alloc(size_t(ntri) * 3 * sizeof(float), s));
            k_rquant_tri<C, T><<<dim3(unsigned((3 * ntri + kB - 1) / kB)), dim3(kB), 0, s>>>(dv.as<T>(), di.as<int32_t>(),
                                                                                         3 * ntri, Q, dt.as<v3<C>>());
            AGWN_HIP(hipGetLastError());
            unsigned const g = unsigned((ntri + kB - 1) / kB);
            k_rleaf_boxes<C><<<dim3(g), dim3(kB), 0, s>>>(dt.as<v3<C>>(), ntri, 0.0, nullptr, nullptr, dc.as<float>());
            AGWN_HIP(hipGetLastError());
            AGWN_TRY(lbvh_build_topology(dc.as<float>(), ntri, topo, s));
            AGWN_HIP(hipMalloc(&tri, size_t(3 * ntri) * sizeof(v3<C>)));
            k_rgather<C><<<dim3(g), dim3(kB), 0, s>>>(dt.as<v3<C>>(), topo.order, ntri, tri);
            AGWN_HIP(hipGetLastError());
            if (ntri > 1)
            {
                dbuf lo, hi, nlo, nhi, fl;
                AGWN_TRY(lo.alloc(size_t(ntri) * 3 * sizeof(double), s));
                AGWN_TRY(hi.alloc(size_t(ntri) * 3 * sizeof(double), s));
                AGWN_TRY(nlo.alloc(size_t(ntri) * 3 * sizeof(double), s));
                AGWN_TRY(nhi.alloc(size_t(ntri) * 3 * sizeof(double), s));
                AGWN_TRY(fl.alloc(size_t(ntri) * sizeof(int32_t), s));
                double const grid_pad = bits > 0 ? 1.0 : 0.0;
                k_rleaf_boxes<C><<<dim3(g), dim3(kB), 0, s>>>(tri, ntri, grid_pad, lo.as<double>(), hi.as<double>(),
                                                              nullptr);
                AGWN_HIP(hipGetLastError());
                AGWN_HIP(hipMalloc(&nodes, size_t(ntri - 1) * sizeof(bvh_node<double>)));
                AGWN_TRY(lbvh_refit<double>(topo, lo.as<double>(), hi.as<double>(), nodes, nlo.as<double>(),
                                            nhi.as<double>(), fl.as<int32_t>(), s));
                AGWN_HIP(hipStreamSynchronize(s));
            }
        }
        AGWN_HIP(hipStreamSynchronize(s));
        return ok;
    }

    int count(void const* pos, int64_t n, int32_t* out, hipStream_t s) const override
    {
        if (n == 0)
            return ok;
        int64_t const blocks = (n + kTB - 1) / kTB;
        AGWN_CHECK(blocks < (int64_t(1) << 31), "too many queries (%lld)", (long long)n);
        k_rcount<C, T><<<dim3(unsigned(blocks)), dim3(kTB), 0, s>>>(static_cast<T const*>(pos), n, Q, nodes, tri, ntri,
                                                                    out);
        AGWN_HIP(hipGetLastError());
        return ok;
    }

    int frac(void const* pos, int64_t n, int32_t const* counts, void* out_frac, void* out_eval,
             hipStream_t s) const override
    {
        if (n == 0)
            return ok;
        int64_t const blocks = (n + kB - 1) / kB;
        AGWN_CHECK(blocks < (int64_t(1) << 31), "too many queries (%lld)", (long long)n);
        constexpr double f_norm = 2.0 / (4.0 * kPi);
        k_rfrac<C, T><<<dim3(unsigned(blocks)), dim3(kB), 0, s>>>(static_cast<T const*>(pos), n, Q, stream, nstream, f_norm,
                                                                  counts, static_cast<T*>(out_frac),
                                                                  static_cast<T*>(out_eval));
        AGWN_HIP(hipGetLastError());
        return ok;
    }
};

template <class T>
int make_robust(int precision, T const* v, int64_t nv, int32_t const* idx, int64_t nidx, T const* seg, int64_t ns,
                std::unique_ptr<robust_impl>& out)
{
    if (precision == AGWN_INT64_GRID)
    {
        auto r = std::make_unique<robust_t<int32_t, T>>();
        AGWN_TRY(r->create(v, nv, idx, nidx, seg, ns, 20));
        out = std::move(r);
    }
    else if (precision == AGWN_INT128_GRID)
    {
        auto r = std::make_unique<robust_t<int64_t, T>>();
        AGWN_TRY(r->create(v, nv, idx, nidx, seg, ns, 50));
        out = std::move(r);
    }
    else if (precision == AGWN_EXACT_FLOAT)
    {
        auto r = std::make_unique<robust_t<double, T>>();
        AGWN_TRY(r->create(v, nv, idx, nidx, seg, ns, 0));
        out = std::move(r);
    }
    else
    {
        set_error("unknown precision policy %d", precision);
        return err_invalid;
    }
    out->scalar_bytes = int(sizeof(T));
    out->precision = precision;
    return ok;
}
} // namespace
} // namespace agwn

struct agwn_robust
{
    std::unique_ptr<agwn::robust_impl> impl;
};

namespace agwn
{
namespace
{
template <class F>
int guarded3(F&& f)
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

int robust_create(int scalar, int precision, void const* verts, int64_t nv, int32_t const* idx, int64_t nidx,
                  void const* seg7, int64_t nseg, agwn_robust** out)
{
    AGWN_CHECK(out != nullptr, "out is null");
    *out = nullptr;
    AGWN_CHECK(scalar == 4 || scalar == 8, "scalar_bytes must be 4 or 8 (got %d)", scalar);
    AGWN_CHECK(nv >= 0 && nidx >= 0 && nseg >= 0, "negative count");
    AGWN_CHECK(nidx % 3 == 0, "indices.size() = %lld is not a multiple of 3", (long long)nidx);
    AGWN_CHECK(nidx == 0 || (verts && idx), "null mesh arrays");
    AGWN_CHECK(nseg == 0 || seg7, "boundary is null with nseg=%lld", (long long)nseg);
    for (int64_t i = 0; i < nidx; ++i)
        AGWN_CHECK(idx[i] >= 0 && idx[i] < nv, "index %lld = %d out of range [0, %lld)", (long long)i, idx[i],
                   (long long)nv);
    AGWN_TRY(ensure_device());
    std::unique_ptr<robust_impl> impl;
    if (scalar == 4)
        AGWN_TRY(make_robust<float>(precision, static_cast<float const*>(verts), nv, idx, nidx,
                                    static_cast<float const*>(seg7), nseg, impl));
    else
        AGWN_TRY(make_robust<double>(precision, static_cast<double const*>(verts), nv, idx, nidx,
                                     static_cast<double const*>(seg7), nseg, impl));
    auto* h = new agwn_robust;
    h->impl = std::move(impl);
    *out = h;
    return ok;
}

enum what
{
    w_eval,
    w_frac,
    w_count
};

int robust_dev(agwn_robust const* r, void const* pos, int64_t n, void* out, what w, hipStream_t s)
{
    AGWN_CHECK(r && r->impl, "robust handle is null");
    AGWN_CHECK(n >= 0, "negative count");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    if (n == 0)
        return ok;
    robust_impl const& R = *r->impl;
    if (w == w_count)
        return R.count(pos, n, static_cast<int32_t*>(out), s);
    if (w == w_frac)
        return R.frac(pos, n, nullptr, out, nullptr, s);
    dbuf cnt;
    AGWN_TRY(cnt.alloc(size_t(n) * sizeof(int32_t), s));
    AGWN_TRY(R.count(pos, n, cnt.as<int32_t>(), s));
    return R.frac(pos, n, cnt.as<int32_t>(), nullptr, out, s);
}

int robust_host(agwn_robust const* r, void const* pos, int64_t n, void* out, what w)
{
    AGWN_CHECK(r && r->impl, "robust handle is null");
    AGWN_CHECK(n >= 0, "negative count");
    AGWN_CHECK(n == 0 || (pos && out), "positions/out are null with n=%lld", (long long)n);
    if (n == 0)
        return ok;
    size_t const sb = size_t(r->impl->scalar_bytes);
    size_t const ob = w == w_count ? sizeof(int32_t) : sb;
    hipStream_t s = nullptr;
    AGWN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct sguard
    {
        hipStream_t s;
        ~sguard() { (void)hipStreamDestroy(s); }
    } sg{s};
    dbuf dp, dout;
    AGWN_TRY(dp.alloc(size_t(n) * 3 * sb, s));
    AGWN_TRY(dout.alloc(size_t(n) * ob, s));
    AGWN_HIP(hipMemcpyAsync(dp.p, pos, size_t(n) * 3 * sb, hipMemcpyHostToDevice, s));
    AGWN_TRY(robust_dev(r, dp.p, n, dout.p, w, s));
    AGWN_HIP(hipMemcpyAsync(out, dout.p, size_t(n) * ob, hipMemcpyDeviceToHost, s));
    AGWN_HIP(hipStreamSynchronize(s));
    return ok;
}
} // namespace
} // namespace agwn

using namespace agwn;

extern "C" {
int agwn_robust_create(int scalar_bytes, int precision, void const* verts, int64_t nv, int32_t const* idx,
                       int64_t nidx, void const* seg7, int64_t nseg, agwn_robust** out)
{
    return guarded3([&] { return robust_create(scalar_bytes, precision, verts, nv, idx, nidx, seg7, nseg, out); });
}
int agwn_robust_destroy(agwn_robust* r)
{
    clear_error();
    delete r;
    return ok;
}
int agwn_robust_info(agwn_robust const* r, int* scalar_bytes, int* precision, int64_t* ntri, int64_t* nseg,
                     double* center3, double* scale)
{
    clear_error();
    AGWN_CHECK(r && r->impl, "robust handle is null");
    robust_impl const& R = *r->impl;
    if (scalar_bytes)
        *scalar_bytes = R.scalar_bytes;
    if (precision)
        *precision = R.precision;
    if (ntri)
        *ntri = R.ntri;
    if (nseg)
        *nseg = R.nseg;
    if (center3)
    {
        center3[0] = R.Q.cx;
        center3[1] = R.Q.cy;
        center3[2] = R.Q.cz;
    }
    if (scale)
        *scale = R.Q.scale;
    return ok;
}
int agwn_robust_eval_batch(agwn_robust const* r, void const* pos, int64_t n, void* out)
{
    return guarded3([&] { return robust_host(r, pos, n, out, w_eval); });
}
int agwn_robust_fractional_batch(agwn_robust const* r, void const* pos, int64_t n, void* out)
{
    return guarded3([&] { return robust_host(r, pos, n, out, w_frac); });
}
int agwn_robust_count_batch(agwn_robust const* r, void const* pos, int64_t n, int32_t* out)
{
    return guarded3([&] { return robust_host(r, pos, n, out, w_count); });
}
int agwn_robust_eval_batch_dev(agwn_robust const* r, void const* pos, int64_t n, void* out, agwn_stream s)
{
    return guarded3([&] { return robust_dev(r, pos, n, out, w_eval, hipStream_t(s)); });
}
int agwn_robust_fractional_batch_dev(agwn_robust const* r, void const* pos, int64_t n, void* out, agwn_stream s)
{
    return guarded3([&] { return robust_dev(r, pos, n, out, w_frac, hipStream_t(s)); });
}
int agwn_robust_count_batch_dev(agwn_robust const* r, void const* pos, int64_t n, int32_t* out, agwn_stream s)
{
    return guarded3([&] { return robust_dev(r, pos, n, out, w_count, hipStream_t(s)); });
}
} // extern "C"
