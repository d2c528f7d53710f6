// Fractional GWN term on gfx950: the boundary sum of gwn_mesh.hh:34-46 / math/common.hh:181-189, batched.
//
// Work decomposition. One lane owns Q query points; a 256-lane workgroup owns 256*Q consecutive queries and walks the
// whole boundary in tiles staged through LDS (one coalesced load per lane, then broadcast reads of the same address
// by every lane).
//
// Pole frame: rows e1 = x1/|x1| (x1 = -x0), e2, e3 orthonormal and right-handed; coordinates (X, y, z) of
// v = endpoint - query. With s = |x1| -- NOT assumed to be 1: a normalised float x0 is off unit length by ~1e-8, and
// the reference's formula is not invariant to |x1| (near the pole the difference reaches 1e-4 rad per pair), so the
// reference's function is evaluated as written, in an exactly equivalent arrangement:
//   num = x1 . (v0 x v1)                       = s (y0 z1 - z0 y1)
//   den = l0 l1 + l1 x1.v0 + l0 x1.v1 + v0.v1  = (l0 + s X0)(l1 + s X1) + (1 - s^2) X0 X1 + y0 y1 + z0 z1
//   l + s X = (X >= 0 ? l + X : (y^2 + z^2) / (l - X)) + (s - 1) X
//   area += w * atan2(num, den)
// The second form has no cancellation when an endpoint lies along the pole (X ~ -l), which the first suffers for
// every segment seen near the pole direction: in FP32 those pairs would carry errors of ~eps * l0 l1 / |den| each,
// and a boundary of 4e5 segments accumulates them past 1e-5.
// FP32 kernels: the pair in FP32, accumulated in FP64 (one convert + one add per term; cheaper than FP32 two-sum
// compensation and more accurate). Coordinates: the rotation is done in FP64, and every pole-frame coordinate of an
// endpoint and of a query is kept as a float pair (hi, lo) with hi + lo = the FP64 value to ~48 bits; v is formed as
// (a_hi - p_hi) + (a_lo - p_lo), so each component is accurate to ~2 eps of itself. Rounding the rotated coordinates
// to one float instead would leave an absolute error of ~eps |a| in v, and a segment at distance l from the query
// would carry a relative error of ~eps |a| / l: 1e-4 rad at l = 1e-3 on a unit-size mesh. A pair with |num| + |den|
// small against the scale of its products (lp0 lp1 + (|y0| + |z0|)(|y1| + |z1|)) is ill-conditioned -- the query
// is near the segment, or the segment passes near the pole direction -- and is re-evaluated in FP64, then scaled
// into FP32 range and finished with the FP32 atan2 (accurate num and den are what matters). So is every pair with
// den <= 0 (there atan2 sits on its branch cut, and the sign of num decides between +pi and -pi).
// FP64 kernel: every pair in FP64 in the same form, with two-sum compensated accumulation.
//
// FP32 chain kernel (boundaries whose weights are small integers -- build_boundary_segments' are +-1): the boundary
// is a stream of points in which consecutive unit segments share endpoints, so each query computes v = P - q and
// l + s X once per point rather than twice per segment; and the atan2 of two consecutive well-conditioned pairs is
// taken once, of the product of their complex numbers (den + i num): with den > 0 each angle lies in (-pi/2, pi/2),
// so the product's argument is the exact sum of the two with no 2 pi wrap. Pairs that go to FP64 enter the product
// as 1.

#include "../../include/agwn.h"
#include "agwn_internal.hpp"
#include "agwn_math.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>
#include <utility>

namespace agwn
{
namespace
{
constexpr int kBlock = 256;
constexpr double kTwoPi = 6.283185307179586476925286766559;

frame make_frame_impl(double x0x, double x0y, double x0z)
{
    double ex = -x0x, ey = -x0y, ez = -x0z;
    double const l = std::sqrt(ex * ex + ey * ey + ez * ez);
    if (l > 0.0)
    {
        ex /= l;
        ey /= l;
        ez /= l;
    }
    else
    {
        ex = 1.0;
        ey = 0.0;
        ez = 0.0;
    }
    // helper axis least aligned with e1
    double hx = 0.0, hy = 0.0, hz = 0.0;
    if (std::fabs(ex) < 0.9)
        hx = 1.0;
    else
        hy = 1.0;
    // e2 = normalize(e1 x h), e3 = e1 x e2  -> rows (e1, e2, e3) are right-handed and orthonormal
    double fx = ey * hz - ez * hy, fy = ez * hx - ex * hz, fz = ex * hy - ey * hx;
    double const lf = std::sqrt(fx * fx + fy * fy + fz * fz);
    fx /= lf;
    fy /= lf;
    fz /= lf;
    double const gx = ey * fz - ez * fy, gy = ez * fx - ex * fz, gz = ex * fy - ey * fx;
    return frame{{ex, ey, ez, fx, fy, fz, gx, gy, gz}};
}

// Rotates and packs the boundary; slots nseg..npad-1 get zero-weight segments far along +e1 (they add exactly 0 and
// are never ill-conditioned), so the FP32 kernel walks whole chunks.
template <class T>
__global__ void __launch_bounds__(kBlock) k_prep_segments(T const* __restrict__ seg7, int64_t nseg, int64_t npad,
                                                          frame R, seg_f32* __restrict__ s32,
                                                          seg_f64* __restrict__ s64)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npad)
        return;
    if (i >= nseg)
    {
        if (s32)
            s32[i] = seg_f32{1.0e6f, 0.0f, 0.0f, 1.0e6f, 0.0f, 0.0f, 0.0f, 0.0f,
                             0.0f,   0.0f, 0.0f, 0.0f,   0.0f, 0.0f, 0.0f, 0.0f};
        s64[i] = seg_f64{1.0e6, 0.0, 0.0, 1.0e6, 0.0, 0.0, 0.0, 0.0};
        return;
    }
    T const* g = seg7 + 7 * i;
    double const ax = double(g[0]), ay = double(g[1]), az = double(g[2]);
    double const bx = double(g[3]), by = double(g[4]), bz = double(g[5]);
    double const w = double(g[6]);
    double const rax = R.r[0] * ax + R.r[1] * ay + R.r[2] * az;
    double const ray = R.r[3] * ax + R.r[4] * ay + R.r[5] * az;
    double const raz = R.r[6] * ax + R.r[7] * ay + R.r[8] * az;
    double const rbx = R.r[0] * bx + R.r[1] * by + R.r[2] * bz;
    double const rby = R.r[3] * bx + R.r[4] * by + R.r[5] * bz;
    double const rbz = R.r[6] * bx + R.r[7] * by + R.r[8] * bz;
    if (s32)
    {
        float const hax = float(rax), hay = float(ray), haz = float(raz);
        float const hbx = float(rbx), hby = float(rby), hbz = float(rbz);
        s32[i] = seg_f32{hax, hay, haz, hbx, hby, hbz, float(w), 0.0f,
                         float(rax - double(hax)), float(ray - double(hay)), float(raz - double(haz)),
                         float(rbx - double(hbx)), float(rby - double(hby)), float(rbz - double(hbz)), 0.0f, 0.0f};
    }
    s64[i] = seg_f64{rax, ray, raz, rbx, rby, rbz, w, 0.0};
}

// |x1| and the exact small quantities derived from it, for the FP64 evaluation
struct pole
{
    double s;   // |x1|
    double sm1; // s - 1
    double c;   // 1 - s^2
};

// One endpoint's l + s X (v = endpoint - query in the pole frame, l = |v|) in FP64 as a fraction lp = Pn / m with a
// positive m and no division: X >= 0: Pn = l + X + (s - 1) X, m = 1; X < 0: l + X = (y^2 + z^2) / (l - X), so
// m = l - X and Pn = y^2 + z^2 + (s - 1) X m.
__device__ __forceinline__ void lplus_f64(double x, double y, double z, double sm1, double& Pn, double& m)
{
    double const yz = __builtin_fma(y, y, z * z);
    double const l = sqrt_f64(__builtin_fma(x, x, yz));
    bool const pos = x >= 0.0;
    double const lx = l + __builtin_fabs(x); // l + x for x >= 0, l - x below: no cancellation either way
    m = pos ? 1.0 : lx;
    Pn = __builtin_fma(sm1 * x, m, pos ? lx : yz);
}

// One pair's (num, den) in FP64, cancellation-free, both multiplied by the positive m0 m1 (the atan2 is unchanged):
// num = s (y0 z1 - z0 y1) m0 m1, den = Pn0 Pn1 + m0 m1 ((1 - s^2) X0 X1 + y0 y1 + z0 z1). v0 = a - p, v1 = b - p in the
// pole frame.
__device__ __forceinline__ void pair_nd_f64(double v0x, double v0y, double v0z, double v1x, double v1y, double v1z,
                                            pole const& P, double& num, double& den)
{
    double P0, m0, P1, m1;
    lplus_f64(v0x, v0y, v0z, P.sm1, P0, m0);
    lplus_f64(v1x, v1y, v1z, P.sm1, P1, m1);
    double const w = m0 * m1;
    num = (P.s * w) * __builtin_fma(v0y, v1z, -(v0z * v1y));
    den = __builtin_fma(P0, P1, w * __builtin_fma(P.c * v0x, v1x, __builtin_fma(v0y, v1y, v0z * v1z)));
}

// One pair in FP64 (the angle).
__device__ __forceinline__ double pair_f64(double v0x, double v0y, double v0z, double v1x, double v1y, double v1z,
                                           pole const& P)
{
    double num, den;
    pair_nd_f64(v0x, v0y, v0z, v1x, v1y, v1z, P, num, den);
    return atan2_f64(num, den);
}

// atan2 for the FP32 kernel: octant reduction to r = min(|y|,|x|) / max(|y|,|x|) in [0, 1] and a degree-15 odd
// minimax polynomial atan(r) ~ r P(r^2) fitted here (Lawson's iteratively reweighted least squares on [0, 1]),
// |error| <= 1.1e-7 evaluated in FP32. Quadrant fix-ups follow C's atan2 (signed zeros included). Callers pass finite
// arguments whose max(|x|, |y|) is normal or zero (the pair range described below, FP64 values scaled near 1,
// atan2_f32_any), so the FLT_MIN floor on the divisor only acts at (0, 0), where r = 0 is exact.
__device__ __forceinline__ float atan2_f32(float y, float x)
{
    constexpr float c0 = 0.9999993443489075f, c1 = -0.33329862356185913f, c2 = 0.19946569204330444f,
                    c3 = -0.13908647000789642f, c4 = 0.09642244875431061f, c5 = -0.055913012474775314f,
                    c6 = 0.021863453090190887f, c7 = -0.004054709803313017f;
    constexpr float half_pi = 1.5707963267948966f, pi = 3.141592653589793f;
    float const ax = __builtin_fabsf(x), ay = __builtin_fabsf(y);
    float const mx = __builtin_fmaxf(ax, ay), mn = __builtin_fminf(ax, ay);
    float const r = mn * __builtin_amdgcn_rcpf(__builtin_fmaxf(mx, 1.17549435e-38f));
    float const t = r * r;
    float p = c7;
    p = __builtin_fmaf(p, t, c6);
    p = __builtin_fmaf(p, t, c5);
    p = __builtin_fmaf(p, t, c4);
    p = __builtin_fmaf(p, t, c3);
    p = __builtin_fmaf(p, t, c2);
    p = __builtin_fmaf(p, t, c1);
    p = __builtin_fmaf(p, t, c0);
    float a = r * p;
    a = ay > ax ? half_pi - a : a;
    a = __builtin_signbit(x) ? pi - a : a;
    return __builtin_copysignf(a, y);
}

// l + s X of one endpoint (v = endpoint - query in the pole frame) without cancellation:
// (X >= 0 ? l + X : (y^2 + z^2) / (l - X)) + (s - 1) X, with l = |v|.
__device__ __forceinline__ float lplus_f32(float x, float yz, float l2, float sm1)
{
    float const l = __builtin_amdgcn_sqrtf(l2);
    float const t = x >= 0.0f ? l + x : yz * __builtin_amdgcn_rcpf(l - x);
    return t + sm1 * x;
}

// The FP32 kernels' ill-conditioned pair: (num, den) formed in FP64 without cancellation (as pair_f64), scaled by a
// power of two into FP32 range and finished with the FP32 atan2. The conditioning problem is in forming num and den,
// not in the arctangent of accurate values: rounding them to FP32 moves the angle by at most ~6e-8 rad, and the
// FP64 library atan2 (several times the cost of the whole FP32 pair) is avoided.
__device__ __forceinline__ float pair_f64_atan_f32(double v0x, double v0y, double v0z, double v1x, double v1y,
                                                   double v1z, pole const& P)
{
    double const yz0 = v0y * v0y + v0z * v0z;
    double const yz1 = v1y * v1y + v1z * v1z;
    double const l0 = sqrt(v0x * v0x + yz0);
    double const l1 = sqrt(v1x * v1x + yz1);
    double const lp0 = (v0x >= 0.0 ? l0 + v0x : yz0 / (l0 - v0x)) + P.sm1 * v0x;
    double const lp1 = (v1x >= 0.0 ? l1 + v1x : yz1 / (l1 - v1x)) + P.sm1 * v1x;
    double const num = P.s * (v0y * v1z - v0z * v1y);
    double const den = lp0 * lp1 + (P.c * v0x * v1x + (v0y * v1y + v0z * v1z));
    double const m = fmax(fabs(num), fabs(den));
    if (!(m > 0.0))
        return atan2_f32(float(num), float(den)); // both zero: the signs decide, as in atan2
    int const e = ilogb(m);
    return atan2_f32(float(ldexp(num, -e)), float(ldexp(den, -e)));
}

// Range of the FP32 pair arithmetic. S = lp0 lp1 + (|y0| + |z0|)(|y1| + |z1|) bounds every product in a pair's num
// and den, and a pair that passes the conditioning test has tau S / sqrt(2) <= |(den, num)| <= 2 S (tau = 1/32). The
// FP32 kernels work in coordinates scaled by the power of two that brings the boundary's extent into [1, 2)
// (boundary_scale_exp), so S is relative to the boundary's size and no result depends on a power-of-two scaling of
// all inputs. A query within 2^26 extents of the centre of the boundary's box then has S <= 2^58 for every pair, so
// the complex product of two pairs stays below 2^118 and finite. A query farther out is "far": its lane runs with an
// infinite conditioning threshold (kFarTau), which sends every one of its pairs to the FP64 path, whose own scaling
// covers any finite value. The test costs a few instructions per query, none per pair. (Below, S is small only for a
// query within ~2^-40 extents of a boundary point, where the fractional term itself varies too fast for FP32
// coordinates to pin it.)
constexpr float kFarTau = 0x1p100f;
__device__ __forceinline__ float lane_tau(float tau, float x, float y, float z, float4 const& far)
{
    float const m = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(x - far.x), __builtin_fabsf(y - far.y)),
                                    __builtin_fabsf(z - far.z));
    return m > far.w ? kFarTau : tau;
}
// atan2_f32 for any finite arguments: below 2^-60, both are scaled by 2^64 first (exact: subnormals become normal)
__device__ __forceinline__ float atan2_f32_any(float y, float x)
{
    if (__builtin_fmaxf(__builtin_fabsf(x), __builtin_fabsf(y)) < 0x1p-60f)
    {
        x *= 0x1p64f;
        y *= 0x1p64f;
    }
    return atan2_f32(y, x);
}
// the conditioning threshold of the lane that owns query i (kFarTau for a far query)
template <class Src>
__device__ __forceinline__ float query_tau(Src const& src, int64_t i, int64_t n, float tau, float4 const& far)
{
    float x = 0.0f, y = 0.0f, z = 0.0f;
    if (i < n)
        query_at(src, i, x, y, z);
    return lane_tau(tau, x, y, z, far);
}

// x unchanged, but opaque to the optimiser: work that depends on it stays where it is written (used to keep the
// FP64 conversions of the rare ill-conditioned branch from being hoisted onto the common path). Emits no instruction.
__device__ __forceinline__ float opaque(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}

template <class Src>
__device__ __forceinline__ void query_f64(Src const& src, int64_t i, int64_t n, frame const& R, double& x, double& y,
                                          double& z)
{
    double px = 0.0, py = 0.0, pz = 0.0;
    if (i < n)
    {
        float fx, fy, fz;
        query_at(src, i, fx, fy, fz);
        px = double(fx);
        py = double(fy);
        pz = double(fz);
    }
    x = R.r[0] * px + R.r[1] * py + R.r[2] * pz;
    y = R.r[3] * px + R.r[4] * py + R.r[5] * pz;
    z = R.r[6] * px + R.r[7] * py + R.r[8] * pz;
}

// A query's pole-frame coordinates as float pairs: h = float(FP64 value), l = float(remainder).
template <class Src>
__device__ __forceinline__ void query_split(Src const& src, int64_t i, int64_t n, frame const& R, float h[3],
                                            float l[3])
{
    double c[3];
    query_f64(src, i, n, R, c[0], c[1], c[2]);
#pragma unroll
    for (int k = 0; k < 3; ++k)
    {
        h[k] = float(c[k]);
        l[k] = float(c[k] - double(h[k]));
    }
}

constexpr int kChunk = 4;          // the boundary is padded to a multiple of this (zero-weight segments)
constexpr int kTile = kBlock;      // segments staged in LDS per pass: one 64-byte segment per lane

template <int Q, class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f32(Src const src, int64_t n, seg_f32 const* __restrict__ seg, seg_f64 const* __restrict__ seg64,
               int64_t nseg_padded, frame R, pole P, float tau, float4 far, int32_t const* __restrict__ counts,
               float* __restrict__ out)
{
    __shared__ seg_f32 tile[kTile];
    int64_t const base = int64_t(blockIdx.x) * (kBlock * Q) + threadIdx.x;
    float ph[Q][3], pl[Q][3], tq[Q];
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        query_split(src, base + int64_t(q) * kBlock, n, R, ph[q], pl[q]);
        tq[q] = query_tau(src, base + int64_t(q) * kBlock, n, tau, far);
        acc[q] = 0.0;
    }
    float const sf = float(P.s), sm1 = float(P.sm1), cf = float(P.c);

    for (int64_t t0 = 0; t0 < nseg_padded; t0 += kTile)
    {
        __syncthreads(); // the previous tile has been consumed by every wave
        if (t0 + threadIdx.x < nseg_padded)
            tile[threadIdx.x] = seg[t0 + threadIdx.x];
        __syncthreads();
        int const cnt = int(nseg_padded - t0 < kTile ? nseg_padded - t0 : kTile);
#pragma unroll 2
        for (int j = 0; j < cnt; ++j)
        {
            seg_f32 const g = tile[j]; // same address in every lane: an LDS broadcast
            double const w = double(g.w);
#pragma unroll
            for (int q = 0; q < Q; ++q)
            {
                float const v0x = (g.ax - ph[q][0]) + (g.lax - pl[q][0]);
                float const v0y = (g.ay - ph[q][1]) + (g.lay - pl[q][1]);
                float const v0z = (g.az - ph[q][2]) + (g.laz - pl[q][2]);
                float const v1x = (g.bx - ph[q][0]) + (g.lbx - pl[q][0]);
                float const v1y = (g.by - ph[q][1]) + (g.lby - pl[q][1]);
                float const v1z = (g.bz - ph[q][2]) + (g.lbz - pl[q][2]);
                float const yz0 = v0y * v0y + v0z * v0z, yz1 = v1y * v1y + v1z * v1z;
                float const lp0 = lplus_f32(v0x, yz0, v0x * v0x + yz0, sm1);
                float const lp1 = lplus_f32(v1x, yz1, v1x * v1x + yz1, sm1);
                float const num = sf * (v0y * v1z - v0z * v1y);
                float const pp = lp0 * lp1;
                float const den = pp + (cf * (v0x * v1x) + (v0y * v1y + v0z * v1z));
                float const scale = pp + (__builtin_fabsf(v0y) + __builtin_fabsf(v0z)) *
                                             (__builtin_fabsf(v1y) + __builtin_fabsf(v1z));
                double a;
                if (__builtin_expect(__builtin_fabsf(num) + __builtin_fabsf(den) < tq[q] * scale || !(den > 0.0f), 0))
                {
                    // ill-conditioned pair: FP64 from the FP64 copies of the segment and the query
                    seg_f64 const h = seg64[t0 + j];
                    double dx, dy, dz;
                    query_f64(src, base + int64_t(q) * kBlock, n, R, dx, dy, dz);
                    a = double(pair_f64_atan_f32(h.ax - dx, h.ay - dy, h.az - dz, h.bx - dx, h.by - dy, h.bz - dz, P));
                }
                else
                {
                    a = double(atan2_f32(num, den));
                }
                acc[q] = fma(w, a, acc[q]);
            }
        }
    }

#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        int64_t const i = base + int64_t(q) * kBlock;
        if (i < n)
        {
            double r = acc[q] / kTwoPi;
            if (counts)
                r += double(counts[i]);
            out[i] = float(r);
        }
    }
}

// Rotates the point stream into the pole frame, two float4 per point: (hi x, y, z, flag) and (lo x, y, z, 0). With
// pexp (a batch of meshes), point i is also scaled by 2^-pexp[i] (its mesh's exponent, exact in double); otherwise
// R already carries the scale.
__global__ void __launch_bounds__(kBlock) k_prep_stream(float4 const* __restrict__ in, int64_t n, frame R,
                                                       int16_t const* __restrict__ pexp, float4* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    float4 const g = in[i];
    double const x = double(g.x), y = double(g.y), z = double(g.z);
    double rx = R.r[0] * x + R.r[1] * y + R.r[2] * z;
    double ry = R.r[3] * x + R.r[4] * y + R.r[5] * z;
    double rz = R.r[6] * x + R.r[7] * y + R.r[8] * z;
    if (pexp)
    {
        double const sc = __longlong_as_double(int64_t(1023 - int(pexp[i])) << 52); // 2^-e, |e| < 1023
        rx *= sc;
        ry *= sc;
        rz *= sc;
    }
    float const hx = float(rx), hy = float(ry), hz = float(rz);
    out[2 * i] = make_float4(hx, hy, hz, g.w);
    out[2 * i + 1] = make_float4(float(rx - double(hx)), float(ry - double(hy)), float(rz - double(hz)), 0.0f);
}

// The chain kernel's work for one block of B lanes x Q queries: queries base + q*B (lane's share), valid below n;
// the point stream pts[0, 2 npts) (pole frame, hi/lo float4 per point) staged S B points at a time through `tile`
// (2 S B float4 of LDS: more points per barrier, fewer resident blocks).
// Per stream point a lane keeps v = P - q (X, y, z), lp = l + s X and |y| + |z|; per unit segment
//   num = s (y0 z1 - z0 y1),  den = lp0 lp1 + (1 - s^2) X0 X1 + (y0 y1 + z0 z1)
// -- the reference's den, rearranged exactly so that no term cancels when an endpoint lies along the pole. A pair is
// ill-conditioned when |num| + |den| < tau (lp0 lp1 + (|y0| + |z0|)(|y1| + |z1|)) (the scale of every product in
// num and den: the query is near the segment, or the segment passes near the pole direction) and goes to FP64 (the
// coordinates hi + lo, exact in double); so do pairs with den <= 0, where atan2 sits on its branch cut and only an
// accurate num gives the sign.
template <int Q, int B, int S = 1, class Src>
__device__ __forceinline__ void chain_block(Src const& src, int64_t base, int64_t n, float4 const* __restrict__ pts,
                                            int64_t npts, frame const& R, pole const& P, float tau,
                                            float4 const& far, int32_t const* __restrict__ counts,
                                            float* __restrict__ out, double* __restrict__ part, float4* tile)
{
    float ph[Q][3], pl[Q][3];                // the query, hi and lo
    float tq[Q];                             // the lane's conditioning threshold (kFarTau for a far query)
    float vx[Q], vy[Q], vz[Q], vp[Q], va[Q]; // the previous stream point: v, l + s X, |y| + |z|
    float pn[Q], pd[Q];                      // the pending pair's (num, den), when `pending`
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        query_split(src, base + int64_t(q) * B, n, R, ph[q], pl[q]);
        tq[q] = query_tau(src, base + int64_t(q) * B, n, tau, far);
        acc[q] = 0.0;
        vx[q] = vy[q] = vz[q] = vp[q] = va[q] = 0.0f;
        pn[q] = 0.0f;
        pd[q] = 1.0f;
    }
    float const sf = float(P.s), sm1 = float(P.sm1), cf = float(P.c);
    bool pending = false; // uniform: the segment structure is the same for every lane
    int64_t const nf = 2 * npts;

    for (int64_t t0 = 0; t0 < npts; t0 += S * B)
    {
        __syncthreads(); // the previous tile has been consumed by every wave
        int64_t const f0 = 2 * t0;
#pragma unroll
        for (int r = 0; r < 2 * S; ++r)
            if (f0 + r * B + threadIdx.x < nf)
                tile[r * B + threadIdx.x] = pts[f0 + r * B + threadIdx.x];
        __syncthreads();
        int const cnt = int(npts - t0 < S * B ? npts - t0 : S * B);
        // one stream point; the loop below takes two per iteration, so that the previous point's registers are
        // renamed rather than copied
        auto step = [&](int const j) __attribute__((always_inline)) {
            float4 const g = tile[2 * j], gl = tile[2 * j + 1]; // same address in every lane: LDS broadcasts
            bool const seg = __builtin_amdgcn_readfirstlane(__float_as_int(g.w)) != 0;
#pragma unroll
            for (int q = 0; q < Q; ++q)
            {
                float const nx = (g.x - ph[q][0]) + (gl.x - pl[q][0]);
                float const ny = (g.y - ph[q][1]) + (gl.y - pl[q][1]);
                float const nz = (g.z - ph[q][2]) + (gl.z - pl[q][2]);
                float const yz = ny * ny + nz * nz;
                float const np = lplus_f32(nx, yz, nx * nx + yz, sm1);
                float const na = __builtin_fabsf(ny) + __builtin_fabsf(nz);
                if (seg)
                {
                    float num = sf * (vy[q] * nz - vz[q] * ny);
                    float const pp = vp[q] * np;
                    float den = pp + (cf * (vx[q] * nx) + (vy[q] * ny + vz[q] * nz));
                    bool const ill = __builtin_fabsf(num) + __builtin_fabsf(den) < tq[q] * (pp + va[q] * na);
                    if (__builtin_expect(ill || !(den > 0.0f), 0))
                    {
                        // the segment's start point: in this tile, or (j == 0) the last point of the previous one --
                        // a flagged point is never the first of a stream, so then t0 > 0
                        float4 h0, l0;
                        if (j > 0)
                        {
                            h0 = tile[2 * j - 2];
                            l0 = tile[2 * j - 1];
                        }
                        else
                        {
                            h0 = pts[f0 - 2];
                            l0 = pts[f0 - 1];
                        }
                        // hi + lo is exact in double
                        auto d = [](float h, float l) { return double(opaque(h)) + double(opaque(l)); };
                        double const qx = d(ph[q][0], pl[q][0]), qy = d(ph[q][1], pl[q][1]);
                        double const qz = d(ph[q][2], pl[q][2]);
                        acc[q] += double(pair_f64_atan_f32(d(h0.x, l0.x) - qx, d(h0.y, l0.y) - qy,
                                                           d(h0.z, l0.z) - qz, d(g.x, gl.x) - qx, d(g.y, gl.y) - qy,
                                                           d(g.z, gl.z) - qz, P));
                        num = 0.0f;
                        den = 1.0f;
                    }
                    if (pending)
                    {
                        float const zr = pd[q] * den - pn[q] * num;
                        float const zi = pd[q] * num + pn[q] * den;
                        acc[q] += double(atan2_f32(zi, zr));
                    }
                    else
                    {
                        pn[q] = num;
                        pd[q] = den;
                    }
                }
                vx[q] = nx;
                vy[q] = ny;
                vz[q] = nz;
                vp[q] = np;
                va[q] = na;
            }
            if (seg)
                pending = !pending;
        };
        int j = 0;
        for (; j + 1 < cnt; j += 2)
        {
            step(j);
            step(j + 1);
        }
        if (j < cnt)
            step(j);
    }

#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        if (pending)
            acc[q] += double(atan2_f32_any(pn[q], pd[q]));
        int64_t const i = base + int64_t(q) * B;
        if (i < n)
        {
            if (part) // a chunk of the stream: the raw sum, added up by k_chain_sum
            {
                part[i] = acc[q];
                continue;
            }
            double r = acc[q] / kTwoPi;
            if (counts)
                r += double(counts[i]);
            out[i] = float(r);
        }
    }
}

template <int Q, class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f32_chain(Src const src, int64_t n, float4 const* __restrict__ pts, int64_t npts, frame R, pole P,
                     float tau, float4 far, int32_t const* __restrict__ counts, float* __restrict__ out)
{
    __shared__ float4 tile[2 * kBlock];
    int64_t const base = int64_t(blockIdx.x) * (kBlock * Q) + threadIdx.x;
    chain_block<Q, kBlock>(src, base, n, pts, npts, R, P, tau, far, counts, out, nullptr, tile);
}

// Launches with too few query tiles to fill the GPU: the stream split into chunks at walk starts (cuts, blockIdx.y
// picks the chunk), partial sums per chunk, added in chunk order by k_chain_sum (deterministic). The cuts are walk
// starts preceded by an even number of segments: there the whole-stream chain has no half-formed pair either (the
// pairing runs across walks), so every chunk forms the same pairs, with the same FP32 arithmetic, as
// k_frac_f32_chain -- only the FP64 partial sums are added in another order.
template <int Q, class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f32_chain_part(Src const src, int64_t n, float4 const* __restrict__ pts, int64_t const* __restrict__ cuts,
                          frame R, pole P, float tau, float4 far, double* __restrict__ part)
{
    __shared__ float4 tile[2 * kBlock];
    int64_t const base = int64_t(blockIdx.x) * (kBlock * Q) + threadIdx.x;
    int64_t const c0 = cuts[blockIdx.y], c1 = cuts[blockIdx.y + 1];
    chain_block<Q, kBlock>(src, base, n, pts + 2 * c0, c1 - c0, R, P, tau, far, nullptr, nullptr,
                           part + int64_t(blockIdx.y) * n, tile);
}
__global__ void __launch_bounds__(kBlock) k_chain_sum(double const* __restrict__ part, int chunks, int64_t n,
                                                      int32_t const* __restrict__ counts, float* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    double a = 0.0;
    for (int c = 0; c < chunks; ++c)
        a += part[int64_t(c) * n + i];
    double r = a / kTwoPi;
    if (counts)
        r += double(counts[i]);
    out[i] = float(r);
}

// Persistent chain kernel over a batch of meshes: workgroups take work items (mesh, first query of a B x Q tile)
// from an atomic counter until none are left; each mesh has its own point stream and query range, and counts (the
// integer term) are added per query. Every workgroup ends at the same test. S B stream points per LDS stage.
template <int Q, int B, int S>
__global__ void __launch_bounds__(B)
    k_frac_f32_chain_batch(float const* __restrict__ pos, batch_frac_desc const* __restrict__ meshes,
                           int32_t const* __restrict__ item_mesh, int64_t const* __restrict__ item_q0, int64_t nitems,
                           unsigned long long* __restrict__ counter, float4 const* __restrict__ pts, frame R, pole P,
                           float tau, int32_t const* __restrict__ counts, float* __restrict__ out)
{
    __shared__ float4 tile[2 * S * B];
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
        batch_frac_desc const D = meshes[item_mesh[it]];
        frame Rm; // the mesh's scaled frame: its stream was rotated and scaled by the same 2^-e (k_prep_stream)
#pragma unroll
        for (int k = 0; k < 9; ++k)
            Rm.r[k] = R.r[k] * D.qscale;
        chain_block<Q, B, S>(src_array<float>{pos}, item_q0[it] + threadIdx.x, D.q_end, pts + 2 * D.pt0, D.npts, Rm, P,
                          tau, D.far, counts, out, nullptr, tile);
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// Far-field decision (frac_decide). A closed loop L of the boundary with a spanning surface D inside the ball (c, rho)
// contributes -Omega_D(q) / 4 pi to the fractional term at a query q outside the ball, provided the line through q
// along x0 misses the ball (then neither the pole cone of the pair formula nor the ray of the integer term crosses D,
// so no integer jump is involved). With f(y) = (y - q) / |y - q|^3, x = c - q, d = |x|, u = x / d and D the fan from c
// (vector area a, area A, first moment M_ij = integral over D of (y - c)_i n_j):
//   -Omega_D(q) = a.u / d^2 + (tr M - 3 u.M.u) / d^3 + R,  |R| <= 3 I / (d - rho)^4 <= 3 rho^2 A / (d - rho)^4
// (Taylor expansion of f about c to first order with the integral remainder; |D^2 f[v, v]| <= 6 |v|^2 / r^4, the
// maximum of |-6 t v + (15 t^2 - 3) x| over t = u.v for unit v; with v = y - c integrated over D, I = the integral of
// |y - c|^2 dA, D's second moment of area about c). D's orientation: it closes the mesh, so its boundary
// runs against the loop and a = -(integral of n dA). Loops whose bound exceeds beta, or that fail the ball tests, are
// summed exactly instead.
struct alignas(16) loop_rec
{
    float cx, cy, cz, rho; // ball in the pole frame; rho grown by 0.1 % (+ 1e-6 |c|) so FP32 tests stay conservative
    float ax, ay, az, k3;  // vector area; k3 = 3 I (numerator of the remainder bound; I <= rho^2 A, see pole_record)
    float m[9];            // first moment, pole frame
    float trm;             // its trace
    float cnorm;           // |c|
    int32_t start, count;  // a loop: its stream entries
    int32_t leaf;          // 1 for a loop
    int32_t skip;          // in the loop tree (depth-first order): the first node after this one's subtree
    float pad;
};
static_assert(sizeof(loop_rec) == 96, "loop_rec layout");

constexpr int kDecideBlock = 256;
constexpr float kFourPiF = 12.566370614359172f;

// The exact sum (radians, as the chain kernel accumulates it) of one closed loop's segments at the query (ph, pl),
// evaluated by every lane of the wave: the loop's points are staged 64 at a time into the wave's own LDS slots (one
// hi/lo pair per lane), then read by all lanes. Same arithmetic as chain_block (FP32 pairs, FP64 for ill-conditioned
// and den <= 0 pairs, paired atan2 of complex products).
__device__ double loop_exact(float4 const* __restrict__ p, int count, float const ph[3], float const pl[3],
                             pole const& P, float tau, float4* st, int lane)
{
    float const sf = float(P.s), sm1 = float(P.sm1), cf = float(P.c);
    double acc = 0.0;
    float vx = 0.0f, vy = 0.0f, vz = 0.0f, vp = 0.0f, va = 0.0f;
    float pn = 0.0f, pd = 1.0f;
    bool pending = false;
    float4 h0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), l0 = h0; // the previous point, for the FP64 fallback
    for (int c0 = 0; c0 < count; c0 += 64)
    {
        int const m = count - c0 < 64 ? count - c0 : 64;
        __builtin_amdgcn_wave_barrier(); // the previous chunk has been read
        if (lane < m)
        {
            st[2 * lane] = p[2 * (c0 + lane)];
            st[2 * lane + 1] = p[2 * (c0 + lane) + 1];
        }
        __builtin_amdgcn_wave_barrier();
        for (int j = 0; j < m; ++j)
        {
            float4 const g = st[2 * j], gl = st[2 * j + 1];
            bool const seg = __builtin_amdgcn_readfirstlane(__float_as_int(g.w)) != 0;
            float const nx = (g.x - ph[0]) + (gl.x - pl[0]);
            float const ny = (g.y - ph[1]) + (gl.y - pl[1]);
            float const nz = (g.z - ph[2]) + (gl.z - pl[2]);
            float const yz = ny * ny + nz * nz;
            float const np = lplus_f32(nx, yz, nx * nx + yz, sm1);
            float const na = __builtin_fabsf(ny) + __builtin_fabsf(nz);
            if (seg)
            {
                float num = sf * (vy * nz - vz * ny);
                float const pp = vp * np;
                float den = pp + (cf * (vx * nx) + (vy * ny + vz * nz));
                bool const ill = __builtin_fabsf(num) + __builtin_fabsf(den) < tau * (pp + va * na);
                if (__builtin_expect(ill || !(den > 0.0f), 0))
                {
                    auto d = [](float h, float l) { return double(opaque(h)) + double(opaque(l)); };
                    double const qx = d(ph[0], pl[0]), qy = d(ph[1], pl[1]), qz = d(ph[2], pl[2]);
                    acc += double(pair_f64_atan_f32(d(h0.x, l0.x) - qx, d(h0.y, l0.y) - qy, d(h0.z, l0.z) - qz,
                                                    d(g.x, gl.x) - qx, d(g.y, gl.y) - qy, d(g.z, gl.z) - qz, P));
                    num = 0.0f;
                    den = 1.0f;
                }
                if (pending)
                    acc += double(atan2_f32(pd * num + pn * den, pd * den - pn * num));
                else
                {
                    pn = num;
                    pd = den;
                }
                pending = !pending;
            }
            vx = nx;
            vy = ny;
            vz = nz;
            vp = np;
            va = na;
            h0 = g;
            l0 = gl;
        }
    }
    if (pending)
        acc += double(atan2_f32_any(pn, pd));
    return acc;
}

// Far-field test of one record (a loop, or a node of the loop tree) at the query: true when it must be summed exactly
// (a node: descended into); otherwise its estimate and bound are added to est and bnd (steradians).
__device__ __forceinline__ bool farfield_near(loop_rec const& r, float const ph[3], float beta4pi, float& est,
                                              float& bnd)
{
    float const dx = r.cx - ph[0], dy = r.cy - ph[1], dz = r.cz - ph[2];
    float const yz = dy * dy + dz * dz;
    float const d = __builtin_amdgcn_sqrtf(dx * dx + yz);
    float const gap = d - r.rho;
    // near when the line along x0 (the pole frame's first axis) meets the ball, or the query is in or close to it
    if (!(yz > r.rho * r.rho) || !(gap > 1e-3f * (d + r.cnorm)))
        return true;
    float const g2 = gap * gap;
    float const b = r.k3 * __builtin_amdgcn_rcpf(g2 * g2);
    if (b > beta4pi)
        return true;
    float const inv = __builtin_amdgcn_rcpf(d);
    float const ux = dx * inv, uy = dy * inv, uz = dz * inv;
    float const ad = r.ax * ux + r.ay * uy + r.az * uz;
    float const mq = ux * (r.m[0] * ux + r.m[1] * uy + r.m[2] * uz) + uy * (r.m[3] * ux + r.m[4] * uy + r.m[5] * uz) +
                     uz * (r.m[6] * ux + r.m[7] * uy + r.m[8] * uz);
    est += (ad + (r.trm - 3.0f * mq) * inv) * (inv * inv);
    bnd += b;
    return false;
}

// An inner node of the loop tree is accepted only when its bound is below beta / 8 (a loop: beta). Accepted high in
// the tree at the full beta, the summed bound grew until more points stayed undecided (with the rho^2 A bound the
// benchmark grid's step took 118.8 ms that way against 112.4 ms at beta / 16). With the second-moment bound the exact
// stage of that step took 27.4 ms at beta / 16, 26.7 at beta / 8, 26.8 at beta / 4; beta itself at 0.001 / 0.002 /
// 0.004: 28.4 / 27.4 / 27.3 ms.
constexpr float kInnerBeta = 1.0f / 8.0f;

// One lane per query (FP32 positions in the caller's frame, counts = integer terms). The wave walks the loop tree in
// depth-first order without a stack: a lane whose far-field test passes at node k takes that node's estimate and
// bound for every loop below it and resumes at skip[k]; a lane that fails goes on to k + 1 (the first child; at a
// loop, it sums the loop exactly -- one exact sum per wave, staged through the wave's LDS slots). The wave visits
// k + 1 when any lane failed at k, else skip[k]: every lane that stopped at an ancestor resumes at that ancestor's
// skip, which is no earlier, so the lanes active at a visited node are exactly those whose resume index is that node.
__global__ void __launch_bounds__(kDecideBlock)
    k_frac_decide(float const* __restrict__ pos, int64_t n, int32_t const* __restrict__ counts,
                  loop_rec const* __restrict__ tree, int nnodes, float4 const* __restrict__ pts, frame R, pole P,
                  float tau, float4 far, float beta4pi, double margin, float* __restrict__ val, int64_t* __restrict__ undecided,
                  unsigned long long* __restrict__ nund)
{
    __shared__ float4 stage[kDecideBlock / 64][128];
    int const lane = int(threadIdx.x & 63);
    float4* st = stage[threadIdx.x >> 6];
    int64_t const i = int64_t(blockIdx.x) * kDecideBlock + threadIdx.x;
    bool const valid = i < n;
    float ph[3], pl[3];
    query_split(src_array<float>{pos}, i, n, R, ph, pl);
    float const tl = query_tau(src_array<float>{pos}, i, n, tau, far);
    double fe = 0.0;              // exact loops (radians)
    float est = 0.0f, bnd = 0.0f; // accepted nodes: -Omega estimate and remainder bound (steradians)
    int resume = valid ? 0 : INT32_MAX;
    int k = 0;
    while (k < nnodes)
    {
        loop_rec const r = tree[k];
        bool split = false;
        if (resume == k)
        {
            split = farfield_near(r, ph, r.leaf ? beta4pi : beta4pi * kInnerBeta, est, bnd);
            resume = split ? k + 1 : r.skip;
        }
        bool const down = __any(split);
        if (r.leaf && down)
        {
            double const sl = loop_exact(pts + 2 * int64_t(r.start), r.count, ph, pl, P, tl, st, lane);
            if (split)
                fe += sl;
        }
        k = down ? k + 1 : r.skip;
    }
    if (valid)
    {
        double const wv = double(counts[i]) + fe / kTwoPi + double(est) / (4.0 * 3.141592653589793);
        double const B = double(bnd) / (4.0 * 3.141592653589793) * 1.01 + margin;
        if (fabs(wv - 0.5) > B)
            val[i] = float(wv);
        else
            undecided[atomicAdd(nund, 1ull)] = i;
    }
}

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    double const bv = s - a;
    double const av = s - bv;
    e = (a - av) + (b - bv);
}

template <class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f64(Src const src, int64_t n, seg_f64 const* __restrict__ seg, int64_t nseg, frame R,
               pole P, int32_t const* __restrict__ counts, double* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    double x = 0.0, y = 0.0, z = 0.0;
    if (i < n)
        query_at(src, i, x, y, z);
    double const px = R.r[0] * x + R.r[1] * y + R.r[2] * z;
    double const py = R.r[3] * x + R.r[4] * y + R.r[5] * z;
    double const pz = R.r[6] * x + R.r[7] * y + R.r[8] * z;
    double sum = 0.0, comp = 0.0;
    auto add = [&](double v) {
        double t, e;
        two_sum(sum, v, t, e);
        sum = t;
        comp += e;
    };
    // Segments two at a time: with equal weights and both denominators positive, each angle lies in (-pi/2, pi/2), so
    // their sum is the argument of the product (den0 + i num0)(den1 + i num1) -- one atan2 for two pairs. The product
    // is formed only when both (num, den) magnitudes lie in [2^-400, 2^400] (no overflow or underflow); otherwise, and
    // for unequal weights or a non-positive denominator, each pair takes its own atan2.
    int64_t s = 0;
    for (; s + 1 < nseg; s += 2)
    {
        seg_f64 const g0 = seg[s], g1 = seg[s + 1];
        double n0, d0, n1, d1;
        pair_nd_f64(g0.ax - px, g0.ay - py, g0.az - pz, g0.bx - px, g0.by - py, g0.bz - pz, P, n0, d0);
        pair_nd_f64(g1.ax - px, g1.ay - py, g1.az - pz, g1.bx - px, g1.by - py, g1.bz - pz, P, n1, d1);
        double const m0 = fmax(fabs(n0), d0), m1 = fmax(fabs(n1), d1);
        constexpr double lo = 0x1p-400, hi = 0x1p400;
        if (g0.w == g1.w && d0 > 0.0 && d1 > 0.0 && m0 >= lo && m0 <= hi && m1 >= lo && m1 <= hi)
            add(g0.w * atan2_f64(fma(n0, d1, n1 * d0), fma(d0, d1, -(n0 * n1))));
        else
        {
            add(g0.w * atan2_f64(n0, d0));
            add(g1.w * atan2_f64(n1, d1));
        }
    }
    if (s < nseg)
    {
        seg_f64 const g = seg[s];
        add(g.w * pair_f64(g.ax - px, g.ay - py, g.az - pz, g.bx - px, g.by - py, g.bz - pz, P));
    }
    if (i < n)
    {
        double r = (sum + comp) / kTwoPi;
        if (counts)
            r += double(counts[i]);
        out[i] = r;
    }
}

// The FP64 chain kernel's stream point in the pole frame: (X, y, z, flag), flag 1: a unit segment ends here.
struct alignas(32) pt64
{
    double x, y, z, w;
};

// Rotates a double stream (x, y, z, flag) into the pole frame.
__global__ void __launch_bounds__(kBlock) k_prep_stream64(pt64 const* __restrict__ in, int64_t n, frame R,
                                                         pt64* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    pt64 const g = in[i];
    out[i] = pt64{R.r[0] * g.x + R.r[1] * g.y + R.r[2] * g.z, R.r[3] * g.x + R.r[4] * g.y + R.r[5] * g.z,
                  R.r[6] * g.x + R.r[7] * g.y + R.r[8] * g.z, g.w};
}

// FP64 chain kernel (boundaries of small integer weights, expanded to unit segments and walked as chains): one lane
// per query, the stream staged in LDS 256 points at a time (one 32-byte load per lane, then broadcast reads). Per
// stream point a lane computes v = P - q, l + s X as the division-free fraction (Pn, m) and c X once, where the
// per-segment kernel computes both endpoints of every segment; per unit segment, (num, den) times m0 m1 (see
// pair_nd_f64). Two consecutive pairs whose den are positive and whose magnitudes lie in [2^-400, 2^400] share one
// atan2 (the argument of the product of their complex numbers is the exact sum of two angles in (-pi/2, pi/2));
// others take their own. Two-sum compensated accumulation.
template <class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f64_chain(Src const src, int64_t n, pt64 const* __restrict__ pts, int64_t npts, frame R, pole P,
                     int32_t const* __restrict__ counts, double* __restrict__ out)
{
    __shared__ pt64 tile[kBlock];
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    double x = 0.0, y = 0.0, z = 0.0;
    if (i < n)
        query_at(src, i, x, y, z);
    double const qx = R.r[0] * x + R.r[1] * y + R.r[2] * z;
    double const qy = R.r[3] * x + R.r[4] * y + R.r[5] * z;
    double const qz = R.r[6] * x + R.r[7] * y + R.r[8] * z;
    double sum = 0.0, comp = 0.0;
    auto add = [&](double v) {
        double s, e;
        two_sum_f64(sum, v, s, e);
        sum = s;
        comp += e;
    };
    double vx = 0.0, vy = 0.0, vz = 0.0, vP = 1.0, vm = 1.0; // the previous stream point
    double pn = 0.0, pd = 1.0;                               // the pending pair
    bool pok = false, pending = false;                       // pending is the same in every lane
    constexpr double kLo = 0x1p-400, kHi = 0x1p400;
    for (int64_t t0 = 0; t0 < npts; t0 += kBlock)
    {
        __syncthreads(); // the previous tile has been consumed by every wave
        if (t0 + threadIdx.x < npts)
            tile[threadIdx.x] = pts[t0 + threadIdx.x];
        __syncthreads();
        int const cnt = int(npts - t0 < kBlock ? npts - t0 : kBlock);
        for (int j = 0; j < cnt; ++j)
        {
            pt64 const g = tile[j];
            bool const seg = __builtin_amdgcn_readfirstlane(int(g.w)) != 0;
            double const nx = g.x - qx, ny = g.y - qy, nz = g.z - qz;
            double nP, nm;
            lplus_f64(nx, ny, nz, P.sm1, nP, nm);
            if (seg)
            {
                double const w = vm * nm;
                double const num = (P.s * w) * __builtin_fma(vy, nz, -(vz * ny));
                double const den =
                    __builtin_fma(vP, nP, w * __builtin_fma(P.c * vx, nx, __builtin_fma(vy, ny, vz * nz)));
                double const mag = __builtin_fmax(__builtin_fabs(num), __builtin_fabs(den));
                bool const ok = den > 0.0 && mag >= kLo && mag <= kHi;
                if (pending)
                {
                    if (ok && pok)
                        add(atan2_f64(__builtin_fma(pn, den, pd * num), __builtin_fma(pd, den, -(pn * num))));
                    else
                    {
                        add(atan2_f64(pn, pd));
                        add(atan2_f64(num, den));
                    }
                }
                else
                {
                    pn = num;
                    pd = den;
                    pok = ok;
                }
                pending = !pending;
            }
            vx = nx;
            vy = ny;
            vz = nz;
            vP = nP;
            vm = nm;
        }
    }
    if (pending)
        add(atan2_f64(pn, pd));
    if (i < n)
    {
        double r = (sum + comp) / kTwoPi;
        if (counts)
            r += double(counts[i]);
        out[i] = r;
    }
}

pole make_pole(double x0x, double x0y, double x0z)
{
    double const s2 = x0x * x0x + x0y * x0y + x0z * x0z;
    double const s = std::sqrt(s2);
    double const sm1 = (s2 - 1.0) / (s + 1.0); // s - 1 without cancellation
    return pole{s, sm1, -sm1 * (s + 1.0)};
}

// ill-conditioning threshold of the FP32 kernel: pairs with |num| + |den| < tau * l0 * l1 go to FP64
constexpr float kTau = 1.0f / 32.0f;
constexpr int kQ = 1;
// below kSplitBelow query tiles the chain kernel splits the stream to reach ~kSplitTiles workgroups
constexpr int64_t kSplitBelow = 1024;
constexpr int64_t kSplitTiles = 2048;
} // namespace

frame make_pole_frame(double x0x, double x0y, double x0z) { return make_frame_impl(x0x, x0y, x0z); }

int boundary_scale_exp(float const* h, int64_t nseg, float4* far)
{
    if (far)
        *far = make_float4(0.0f, 0.0f, 0.0f, HUGE_VALF);
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int64_t i = 0; i < nseg; ++i)
        for (int k = 0; k < 6; ++k)
        {
            double const v = double(h[7 * i + k]);
            lo[k % 3] = std::min(lo[k % 3], v);
            hi[k % 3] = std::max(hi[k % 3], v);
        }
    double ext = 0.0;
    for (int a = 0; a < 3; ++a)
        if (hi[a] >= lo[a])
            ext = std::max(ext, hi[a] - lo[a]);
    if (!(ext > 0.0) || !std::isfinite(ext))
        return 0;
    int const e = std::ilogb(ext);
    if (far)
        *far = make_float4(float(0.5 * (lo[0] + hi[0])), float(0.5 * (lo[1] + hi[1])), float(0.5 * (lo[2] + hi[2])),
                           float(std::ldexp(1.0, std::min(26 + e, 127))));
    return e;
}

frame scale_frame(frame const& R, int e)
{
    frame S = R;
    double const sc = std::ldexp(1.0, -e);
    for (double& v : S.r)
        v *= sc;
    return S;
}

// Unit-weight expansion and chain ordering of a host boundary into a point stream (x, y, z, flag). Weights must be
// integers of magnitude <= 16 (0 drops the segment); otherwise returns false. Greedy walks in input order: start at
// the first unused unit segment, follow any unused segment leaving the current end point (exact coordinate match),
// stop when there is none. Every unit segment appears exactly once, as a flag-1 point after its start point.
// emit(p, flag) appends the stream point of coordinates p[0..3).
template <class T, class Emit>
bool chain_walk(T const* h, int64_t nseg, int64_t& nunit, Emit&& emit)
{
    nunit = 0;
    int64_t total = 0;
    for (int64_t i = 0; i < nseg; ++i)
    {
        T const w = h[7 * i + 6];
        if (!(w == std::rint(w)) || std::fabs(w) > T(16))
            return false; // general weights: the per-segment kernel
        total += int64_t(std::fabs(w));
    }
    using word = std::conditional_t<sizeof(T) == 4, uint32_t, uint64_t>;
    struct key
    {
        word b[3];
        bool operator<(key const& o) const
        {
            return b[0] != o.b[0] ? b[0] < o.b[0] : (b[1] != o.b[1] ? b[1] < o.b[1] : b[2] < o.b[2]);
        }
        bool operator==(key const& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
    };
    auto key_of = [](T const* p) {
        key k;
        std::memcpy(k.b, p, sizeof(k.b));
        return k;
    };
    // unit edges (from, to): pointers to the endpoint coordinates
    std::vector<T const*> from, to;
    from.reserve(size_t(total));
    to.reserve(size_t(total));
    for (int64_t i = 0; i < nseg; ++i)
    {
        T const* g = h + 7 * i;
        int const w = int(g[6]);
        for (int c = 0; c < (w < 0 ? -w : w); ++c)
        {
            from.push_back(w > 0 ? g : g + 3);
            to.push_back(w > 0 ? g + 3 : g);
        }
    }
    int64_t const E = int64_t(from.size());
    // edges sorted by their start point; per start point, a cursor over its not yet used edges
    std::vector<int64_t> order(static_cast<size_t>(E));
    for (int64_t e = 0; e < E; ++e)
        order[size_t(e)] = e;
    std::vector<key> fk(static_cast<size_t>(E));
    for (int64_t e = 0; e < E; ++e)
        fk[size_t(e)] = key_of(from[size_t(e)]);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return fk[size_t(a)] < fk[size_t(b)]; });
    std::vector<key> sorted_keys(static_cast<size_t>(E));
    for (int64_t r = 0; r < E; ++r)
        sorted_keys[size_t(r)] = fk[size_t(order[size_t(r)])];
    std::vector<int64_t> cursor(static_cast<size_t>(E)); // valid at the first rank of each key group
    for (int64_t r = 0; r < E; ++r)
        cursor[size_t(r)] = r;
    std::vector<char> used(static_cast<size_t>(E), 0);
    auto next_from = [&](T const* p) -> int64_t {
        key const k = key_of(p);
        auto it = std::lower_bound(sorted_keys.begin(), sorted_keys.end(), k);
        if (it == sorted_keys.end() || !(*it == k))
            return -1;
        int64_t const g = int64_t(it - sorted_keys.begin());
        int64_t& c = cursor[size_t(g)];
        while (c < E && sorted_keys[size_t(c)] == k && used[size_t(order[size_t(c)])])
            ++c;
        if (c < E && sorted_keys[size_t(c)] == k)
            return order[size_t(c)];
        return -1;
    };
    for (int64_t e0 = 0; e0 < E; ++e0)
    {
        if (used[size_t(e0)])
            continue;
        emit(from[size_t(e0)], 0);
        int64_t e = e0;
        while (e >= 0)
        {
            used[size_t(e)] = 1;
            T const* b = to[size_t(e)];
            emit(b, 1);
            e = next_from(b);
        }
    }
    nunit = E;
    return true;
}

bool chain_stream(float const* h, int64_t nseg, std::vector<float4>& stream_pts, int64_t& nunit)
{
    stream_pts.clear();
    stream_pts.reserve(size_t(nseg + nseg / 8 + 16));
    return chain_walk(h, nseg, nunit, [&](float const* p, int flag) {
        stream_pts.push_back(make_float4(p[0], p[1], p[2], float(flag)));
    });
}

// Far-field data of one closed loop (points p[0..n), p[n-1] = p[0]): centroid of the distinct points, radius, vector
// area, fan area and first moment of the fan from the centroid (frac_plan::loops).
loop_host loop_data(float4 const* p, int64_t n)
{
    loop_host L{};
    int64_t const m = n - 1; // distinct points
    for (int64_t i = 0; i < m; ++i)
    {
        L.c[0] += double(p[i].x);
        L.c[1] += double(p[i].y);
        L.c[2] += double(p[i].z);
    }
    for (double& v : L.c)
        v /= double(m);
    for (int64_t i = 0; i < m; ++i)
    {
        double const u[3] = {double(p[i].x) - L.c[0], double(p[i].y) - L.c[1], double(p[i].z) - L.c[2]};
        double const w[3] = {double(p[i + 1].x) - L.c[0], double(p[i + 1].y) - L.c[1], double(p[i + 1].z) - L.c[2]};
        double const t[3] = {0.5 * (u[1] * w[2] - u[2] * w[1]), 0.5 * (u[2] * w[0] - u[0] * w[2]),
                             0.5 * (u[0] * w[1] - u[1] * w[0])}; // n dA of the fan triangle (c, p_i, p_i+1)
        L.rho = std::max(L.rho, std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
        double const at = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        L.area += at;
        // over the triangle (0, u, w): the integral of |y|^2 dA = at (|u|^2 + |w|^2 + u.w) / 6, of y dA = at (u + w) / 3
        L.inertia += at * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + w[0] * w[0] + w[1] * w[1] + w[2] * w[2] +
                           u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) / 6.0;
        for (int r = 0; r < 3; ++r)
        {
            L.a[r] += t[r];
            double const g = (u[r] + w[r]) / 3.0; // triangle centroid - c
            L.s[r] += at * g;
            for (int k = 0; k < 3; ++k)
                L.m[3 * r + k] += g * t[k];
        }
    }
    return L;
}

// A walk longer than kArc segments is split into arcs of kArc segments, each closed by a chord back to its first
// point, plus the skeleton loop through the split points, whose chords run the other way: the chords cancel, so the
// pieces sum to the walk exactly, and every arc is compact (small radius and fan area: a small far-field bound) while
// the skeleton has few points. Long loops would otherwise be summed exactly for most queries.
constexpr int64_t kArc = 16; // the benchmark grid's exact stage: 44.2 ms at 32 and at 8, 42.3 ms at 16, 50.0 at 64

// The stream's walks as closed loops, split as above, into a decision stream of their own (x, y, z, flag: 0 starts a
// loop) with their far-field data; false if some walk is open.
bool loops_of_stream(std::vector<float4> const& p, std::vector<loop_host>& loops, std::vector<float4>& dec)
{
    loops.clear();
    dec.clear();
    std::vector<float4> piece;
    auto emit = [&]() {
        loop_host L = loop_data(piece.data(), int64_t(piece.size()));
        L.start = int64_t(dec.size());
        L.count = int64_t(piece.size());
        for (size_t k = 0; k < piece.size(); ++k)
            dec.push_back(make_float4(piece[k].x, piece[k].y, piece[k].z, k == 0 ? 0.0f : 1.0f));
        loops.push_back(L);
    };
    int64_t const n = int64_t(p.size());
    for (int64_t s0 = 0; s0 < n;)
    {
        int64_t e = s0 + 1;
        while (e < n && p[size_t(e)].w != 0.0f)
            ++e;
        float4 const& a = p[size_t(s0)];
        float4 const& b = p[size_t(e - 1)];
        if (e - s0 < 2 || a.x != b.x || a.y != b.y || a.z != b.z)
            return false;
        int64_t const m = e - 1 - s0; // segments
        if (m <= kArc)
        {
            piece.assign(p.begin() + s0, p.begin() + e);
            emit();
        }
        else
        {
            std::vector<int64_t> ks;
            for (int64_t k = 0; k < m; k += kArc)
                ks.push_back(k);
            ks.push_back(m);
            for (size_t j = 0; j + 1 < ks.size(); ++j)
            {
                piece.assign(p.begin() + s0 + ks[j], p.begin() + s0 + ks[j + 1] + 1);
                piece.push_back(p[size_t(s0 + ks[j])]); // chord back to the arc's first point
                emit();
            }
            piece.clear();
            for (int64_t k : ks)
                piece.push_back(p[size_t(s0 + k)]);
            emit(); // skeleton: first point, split points, first point again
        }
        s0 = e;
    }
    return true;
}

// The loop tree of the decision: the loops in the Morton order of their centres (10 bits per axis over the box of the
// centres), halved recursively. A node's ball is centred where the smallest ball holding its children's balls is, with
// the radius of the farthest loop point below it (no larger than that ball's: every point lies in a child's ball). It
// holds every fan below, as each fan lies in the convex hull of its loop's points. Its vector area and fan area are the
// children's sums and its first moment, about its centre, sum_child M_child + (c_child - c) (x) a_child. The far-field
// expansion and bound hold for any surface inside the ball, so a node is tested and estimated like one loop. Nodes in
// depth-first order, the root first, each with the index just past its subtree (skip). pts: the decision stream (the
// loops' points).
struct tree_node
{
    loop_host L;
    int32_t loop; // >= 0: a leaf, the loop's index
    int32_t skip; // the first node after this one's subtree
};

void loop_tree(std::vector<loop_host> const& loops, std::vector<float4> const& pts, std::vector<tree_node>& nodes)
{
    size_t const n = loops.size();
    nodes.clear();
    if (n == 0)
        return;
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (loop_host const& L : loops)
        for (int a = 0; a < 3; ++a)
        {
            lo[a] = std::min(lo[a], L.c[a]);
            hi[a] = std::max(hi[a], L.c[a]);
        }
    auto spread = [](uint64_t v) {
        uint64_t x = v & 0x3FFu;
        x = (x | (x << 16)) & 0x030000FFull;
        x = (x | (x << 8)) & 0x0300F00Full;
        x = (x | (x << 4)) & 0x030C30C3ull;
        x = (x | (x << 2)) & 0x09249249ull;
        return x;
    };
    std::vector<std::pair<uint64_t, int32_t>> key(n);
    for (size_t k = 0; k < n; ++k)
    {
        uint64_t code = 0;
        for (int a = 0; a < 3; ++a)
        {
            double const ext = hi[a] - lo[a];
            double const f = ext > 0.0 ? (loops[k].c[a] - lo[a]) / ext : 0.0;
            code |= spread(uint64_t(std::min(1023.0, std::max(0.0, f * 1024.0)))) << a;
        }
        key[k] = {code, int32_t(k)};
    }
    std::sort(key.begin(), key.end());
    nodes.reserve(2 * n - 1);
    // builds the node of sorted range [b, e) at the back of `nodes`; returns its index
    auto build = [&](auto&& self, size_t b, size_t e) -> int32_t {
        int32_t const me = int32_t(nodes.size());
        nodes.push_back(tree_node{});
        if (e - b == 1)
        {
            nodes[size_t(me)].L = loops[size_t(key[b].second)];
            nodes[size_t(me)].loop = key[b].second;
            nodes[size_t(me)].skip = me + 1;
            return me;
        }
        size_t const mid = b + (e - b) / 2;
        int32_t const l = self(self, b, mid);
        int32_t const r = self(self, mid, e);
        loop_host const& A = nodes[size_t(l)].L;
        loop_host const& B = nodes[size_t(r)].L;
        loop_host C{};
        double const dv[3] = {B.c[0] - A.c[0], B.c[1] - A.c[1], B.c[2] - A.c[2]};
        double const D = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        if (D + B.rho <= A.rho) // A's ball holds B's
        {
            std::copy(A.c, A.c + 3, C.c);
            C.rho = A.rho;
        }
        else if (D + A.rho <= B.rho)
        {
            std::copy(B.c, B.c + 3, C.c);
            C.rho = B.rho;
        }
        else
        {
            double const t = (D + B.rho - A.rho) / (2.0 * D);
            for (int k = 0; k < 3; ++k)
                C.c[k] = A.c[k] + t * dv[k];
            C.rho = 0.5 * (D + A.rho + B.rho);
        }
        double slack = 0.0; // the magnitudes summed into the second moment (its rounding is a tiny fraction of them)
        for (loop_host const* K : {&A, &B}) // guard the rounding of the centre
        {
            double const dc[3] = {K->c[0] - C.c[0], K->c[1] - C.c[1], K->c[2] - C.c[2]};
            double const dc2 = dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2];
            double const ds = dc[0] * K->s[0] + dc[1] * K->s[1] + dc[2] * K->s[2];
            C.rho = std::max(C.rho, std::sqrt(dc2) + K->rho);
            C.area += K->area;
            // the child's fans about the new centre: I + 2 dc.s + |dc|^2 area, s + area dc
            C.inertia += K->inertia + 2.0 * ds + dc2 * K->area;
            slack += K->inertia + 2.0 * std::fabs(ds) + dc2 * K->area;
            for (int a = 0; a < 3; ++a)
            {
                C.a[a] += K->a[a];
                C.s[a] += K->s[a] + K->area * dc[a];
                for (int j = 0; j < 3; ++j)
                    C.m[3 * a + j] += K->m[3 * a + j] + dc[a] * K->a[j];
            }
        }
        C.inertia = std::max(C.inertia, 0.0) + 1e-12 * slack;
        double far2 = 0.0; // the farthest loop point below from the centre (squared)
        for (size_t k = b; k < e; ++k)
        {
            loop_host const& L = loops[size_t(key[k].second)];
            for (int64_t j = L.start; j < L.start + L.count; ++j)
            {
                float4 const& q = pts[size_t(j)];
                double const u[3] = {double(q.x) - C.c[0], double(q.y) - C.c[1], double(q.z) - C.c[2]};
                far2 = std::max(far2, u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            }
        }
        C.rho = std::min(C.rho, std::sqrt(far2) * (1.0 + 1e-12));
        nodes[size_t(me)].L = C;
        nodes[size_t(me)].loop = -1;
        nodes[size_t(me)].skip = int32_t(nodes.size());
        return me;
    };
    build(build, 0, n);
}

int frac_plan_build(float const* h, int64_t nseg, frac_plan& plan, hipStream_t stream)
{
    plan.dec_state.reset();
    plan.chain = false;
    plan.npts = plan.nunit = 0;
    plan.loops.clear();
    plan.loops_ok = false;
    plan.dec_npts = 0;
    plan.dec_host.clear();
    plan.walk_starts.clear();
    plan.sexp = boundary_scale_exp(h, nseg, &plan.far);
    std::vector<float4> stream_pts;
    int64_t nunit = 0;
    if (!chain_stream(h, nseg, stream_pts, nunit))
        return ok;
    plan.npts = int64_t(stream_pts.size());
    plan.nunit = nunit;
    plan.walk_starts.clear();
    int64_t segs = 0; // segments before entry i
    for (int64_t i = 0; i < plan.npts; ++i)
    {
        if (stream_pts[size_t(i)].w == 0.0f)
        {
            if (segs % 2 == 0)
                plan.walk_starts.push_back(i);
        }
        else
            ++segs;
    }
    std::vector<float4> dec;
    plan.loops_ok = loops_of_stream(stream_pts, plan.loops, dec);
    if (!plan.loops_ok)
        plan.loops.clear();
    else
    {
        plan.dec_npts = int64_t(dec.size());
        plan.dec_host = dec;
        AGWN_TRY(plan.dec.alloc(std::max<size_t>(dec.size(), 1) * sizeof(float4), stream));
        if (!dec.empty())
            AGWN_HIP(hipMemcpyAsync(plan.dec.p, plan.dec_host.data(), dec.size() * sizeof(float4),
                                    hipMemcpyHostToDevice, stream));
    }
    AGWN_TRY(plan.buf.alloc(std::max<size_t>(stream_pts.size(), 1) * sizeof(float4), stream));
    if (!stream_pts.empty())
        AGWN_HIP(hipMemcpyAsync(plan.buf.p, stream_pts.data(), stream_pts.size() * sizeof(float4),
                                hipMemcpyHostToDevice, stream));
    AGWN_HIP(hipStreamSynchronize(stream)); // the host vector goes out of scope
    plan.chain = true;
    return ok;
}

// flag = 1 when the word arrays differ anywhere
__global__ void __launch_bounds__(kBlock) k_words_differ(uint32_t const* __restrict__ a, uint32_t const* __restrict__ b,
                                                         int64_t n, int* __restrict__ flag)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n && a[i] != b[i])
        *flag = 1;
}

int cached_plan(plan_cache& c, float const* seg7, int64_t nseg, hipStream_t s, std::shared_ptr<frac_plan const>& out)
{
    AGWN_CHECK(nseg >= 0 && (nseg == 0 || seg7 != nullptr), "boundary is null with nseg=%lld", (long long)nseg);
    int dev = 0;
    AGWN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.plan && c.nseg == nseg && c.dev == dev)
    {
        int differ = 0;
        if (nseg > 0)
        {
            int64_t const words = nseg * 7;
            AGWN_CHECK((words + kBlock - 1) / kBlock < (int64_t(1) << 31), "boundary too large (%lld segments)",
                       (long long)nseg);
            dbuf flag;
            AGWN_TRY(flag.alloc(sizeof(int), s));
            AGWN_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
            k_words_differ<<<dim3(unsigned((words + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(
                reinterpret_cast<uint32_t const*>(c.seg), reinterpret_cast<uint32_t const*>(seg7), words,
                flag.as<int>());
            AGWN_HIP(hipGetLastError());
            AGWN_HIP(hipMemcpyAsync(&differ, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
            AGWN_HIP(hipStreamSynchronize(s));
        }
        if (!differ)
        {
            out = c.plan;
            return ok;
        }
    }
    std::vector<float> h(size_t(nseg) * 7);
    if (nseg > 0)
    {
        AGWN_HIP(hipMemcpyAsync(h.data(), seg7, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        AGWN_HIP(hipStreamSynchronize(s));
    }
    auto p = std::make_shared<frac_plan>();
    // on the null stream: the plan's buffers are freed on the stream they were allocated on, and the caller's stream
    // may be gone by then
    AGWN_TRY(frac_plan_build(h.data(), nseg, *p, nullptr));
    float* copy = nullptr;
    if (nseg > 0)
    {
        AGWN_HIP(hipMalloc(&copy, h.size() * sizeof(float)));
        hipError_t const e = hipMemcpy(copy, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess)
        {
            (void)hipFree(copy);
            AGWN_HIP(e);
        }
    }
    if (c.seg)
        (void)hipFree(c.seg);
    c.seg = copy;
    c.nseg = nseg;
    c.dev = dev;
    c.plan = p;
    out = p;
    return ok;
}

// The decision margin covers the FP32 evaluations' own errors: the decision's exact loop sums and the plain path's full
// sum, whose label the decision must reproduce. It is empirical, not a derived bound: the FP32 full sum differs from
// the FP64 one by at most 3.8e-6 over 1e7 queries of a 429,869-segment boundary (half of them 1e-3 off the surface)
// and 3.4e-6 on the benchmark's torus, and a worst-case bound per pair summed over ~3e4 pairs would exceed 0.5. Points
// within the margin (plus the far-field bound) of 0.5 go to the full sum, whose arithmetic is the plain path's.
constexpr double kDecideMargin = 1e-4;

// The decision's tables for one pole x0: the loop tree's records in the pole frame (device) and the decision stream
// rotated into that frame (device). They depend on the boundary and x0 only, so a plan keeps them for up to
// kDecidePoles poles (further poles build theirs per call).
struct decide_tables
{
    double x0[3] = {0.0, 0.0, 0.0};
    loop_rec* tab = nullptr;
    int ntab = 0;
    float4* p32 = nullptr;
    decide_tables() = default;
    decide_tables(decide_tables const&) = delete;
    decide_tables& operator=(decide_tables const&) = delete;
    ~decide_tables()
    {
        if (tab)
            (void)hipFree(tab);
        if (p32)
            (void)hipFree(p32);
    }
};
struct decide_state
{
    std::vector<tree_node> nodes; // the loop tree (no pole involved)
    std::vector<std::unique_ptr<decide_tables>> poles;
};
constexpr size_t kDecidePoles = 8;

// a record in the pole frame R (orthonormal), in coordinates scaled by sc (a power of two): lengths times sc, the
// vector area times sc^2, the first moment times sc^3 and the bound's numerator 3 rho^2 A times sc^4, so every test
// and estimate of the decision is the unscaled one; its ball grown by 0.1 % (+ 1e-6 |c|) and its bound numerator by
// 0.1 % so that the FP32 tests stay conservative
loop_rec pole_record(frame const& R, double sc, loop_host const& L, int64_t start, int64_t count)
{
    loop_rec r{};
    double c[3], a[3], t[9], mr[9];
    for (int i = 0; i < 3; ++i)
    {
        c[i] = R.r[3 * i] * L.c[0] + R.r[3 * i + 1] * L.c[1] + R.r[3 * i + 2] * L.c[2];
        a[i] = R.r[3 * i] * L.a[0] + R.r[3 * i + 1] * L.a[1] + R.r[3 * i + 2] * L.a[2];
    }
    for (int i = 0; i < 3; ++i) // t = R M
        for (int j = 0; j < 3; ++j)
            t[3 * i + j] = R.r[3 * i] * L.m[j] + R.r[3 * i + 1] * L.m[3 + j] + R.r[3 * i + 2] * L.m[6 + j];
    for (int i = 0; i < 3; ++i) // mr = t R^T
        for (int j = 0; j < 3; ++j)
            mr[3 * i + j] = t[3 * i] * R.r[3 * j] + t[3 * i + 1] * R.r[3 * j + 1] + t[3 * i + 2] * R.r[3 * j + 2];
    double const cn = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double const rho = L.rho * 1.001 + 1e-6 * cn;
    double const sc2 = sc * sc, sc3 = sc2 * sc;
    r.cx = float(c[0] * sc);
    r.cy = float(c[1] * sc);
    r.cz = float(c[2] * sc);
    r.rho = float(rho * sc);
    r.ax = float(a[0] * sc2);
    r.ay = float(a[1] * sc2);
    r.az = float(a[2] * sc2);
    // remainder bound numerator 3 I: the second-order Taylor term of the kernel at y is at most 3 |y - c|^2 / gap^4,
    // integrated over the fans (I <= rho^2 area: the smaller of the two, both valid)
    double const inertia = std::min(L.inertia, L.rho * L.rho * L.area);
    r.k3 = float(3.0 * inertia * sc2 * sc2 * 1.001);
    for (int i = 0; i < 9; ++i)
        r.m[i] = float(mr[i] * sc3);
    r.trm = float((mr[0] + mr[4] + mr[8]) * sc3);
    r.cnorm = float(cn * sc);
    r.start = int32_t(start);
    r.count = int32_t(count);
    return r;
}

// Builds the tables of pole x0 (frame R) on stream s and waits for them: once kept by the plan they are read by other
// streams' calls.
int build_decide_tables(frac_plan const& plan, std::vector<tree_node> const& nodes, double const x0[3],
                        frame const& R, hipStream_t s, decide_tables& t)
{
    double const sc = std::ldexp(1.0, -plan.sexp);
    frame const Rs = scale_frame(R, plan.sexp);
    for (int a = 0; a < 3; ++a)
        t.x0[a] = x0[a];
    std::vector<loop_rec> tab(nodes.size());
    for (size_t k = 0; k < nodes.size(); ++k)
    {
        tree_node const& N = nodes[k];
        if (N.loop >= 0)
        {
            loop_host const& L = plan.loops[size_t(N.loop)];
            tab[k] = pole_record(R, sc, L, L.start, L.count);
            tab[k].leaf = 1;
        }
        else
        {
            tab[k] = pole_record(R, sc, N.L, 0, 0);
            tab[k].leaf = 0;
        }
        tab[k].skip = N.skip;
    }
    t.ntab = int(tab.size());
    AGWN_HIP(hipMalloc(&t.tab, std::max<size_t>(tab.size(), 1) * sizeof(loop_rec)));
    if (!tab.empty())
        AGWN_HIP(hipMemcpyAsync(t.tab, tab.data(), tab.size() * sizeof(loop_rec), hipMemcpyHostToDevice, s));
    AGWN_HIP(hipMalloc(&t.p32, size_t(std::max<int64_t>(plan.dec_npts, 1)) * 2 * sizeof(float4)));
    if (plan.dec_npts > 0)
    {
        k_prep_stream<<<dim3(unsigned((plan.dec_npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(
            plan.dec.as<float4 const>(), plan.dec_npts, Rs, nullptr, t.p32);
        AGWN_HIP(hipGetLastError());
    }
    AGWN_HIP(hipStreamSynchronize(s)); // the host table goes out of scope; the tables may be shared from now on
    return ok;
}

int frac_decide(float const* pos, int64_t m, int32_t const* counts, frac_plan const& plan, double const x0[3],
                float beta, float* val, int64_t* undecided, int64_t& nund, hipStream_t s, arena* ws)
{
    nund = 0;
    if (m == 0)
        return ok;
    AGWN_CHECK(beta > 0.0f, "beta must be positive");
    AGWN_CHECK(plan.chain && plan.loops_ok, "the far-field decision needs a boundary of closed unit-weight loops");
    AGWN_CHECK(plan.loops.size() < size_t(INT32_MAX) && plan.dec_npts < INT32_MAX, "boundary too large for the decision");
    frame const R = make_frame_impl(x0[0], x0[1], x0[2]);
    pole const P = make_pole(x0[0], x0[1], x0[2]);
    decide_tables const* T = nullptr;
    std::unique_ptr<decide_tables> own; // a pole beyond the kept ones: this call's tables
    {
        std::lock_guard<std::mutex> lk(plan.dec_mu);
        if (!plan.dec_state)
        {
            auto st = std::make_shared<decide_state>();
            loop_tree(plan.loops, plan.dec_host, st->nodes);
            plan.dec_state = st;
        }
        auto* st = static_cast<decide_state*>(plan.dec_state.get());
        for (auto const& e : st->poles)
            if (e->x0[0] == x0[0] && e->x0[1] == x0[1] && e->x0[2] == x0[2])
                T = e.get();
        if (!T)
        {
            auto e = std::make_unique<decide_tables>();
            AGWN_TRY(build_decide_tables(plan, st->nodes, x0, R, s, *e));
            T = e.get();
            if (st->poles.size() < kDecidePoles)
                st->poles.push_back(std::move(e));
            else
                own = std::move(e);
        }
    }
    dbuf dn_own;
    unsigned long long* dn = nullptr;
    if (ws)
        AGWN_TRY(ws->take_n(1, &dn));
    else
    {
        AGWN_TRY(dn_own.alloc(sizeof(unsigned long long), s));
        dn = dn_own.as<unsigned long long>();
    }
    AGWN_HIP(hipMemsetAsync(dn, 0, sizeof(unsigned long long), s));
    int64_t const blocks = (m + kDecideBlock - 1) / kDecideBlock;
    AGWN_CHECK(blocks < (int64_t(1) << 31), "frac_decide: too many queries (%lld)", (long long)m);
    k_frac_decide<<<dim3(unsigned(blocks)), dim3(kDecideBlock), 0, s>>>(
        pos, m, counts, T->tab, T->ntab, T->p32, scale_frame(R, plan.sexp), P, kTau, plan.far, beta * kFourPiF,
        kDecideMargin, val, undecided, dn);
    AGWN_HIP(hipGetLastError());
    unsigned long long h = 0;
    AGWN_HIP(hipMemcpyAsync(&h, dn, sizeof(h), hipMemcpyDeviceToHost, s));
    AGWN_HIP(hipStreamSynchronize(s)); // nund is needed now (and `own` may be freed)
    nund = int64_t(h);
    return ok;
}

template <class Src>
int frac_f32_planned(Src const& pos, int64_t n, frac_plan const& plan, float const* seg7, int64_t nseg,
                     double const x0[3], int32_t const* counts, float* out, hipStream_t stream, bool split_small,
                     arena* ws)
{
    if (n == 0)
        return ok;
    frame const R = scale_frame(make_frame_impl(x0[0], x0[1], x0[2]), plan.sexp);
    pole const P = make_pole(x0[0], x0[1], x0[2]);
    int64_t const per_block = int64_t(kBlock) * kQ;
    int64_t const blocks = (n + per_block - 1) / per_block;
    AGWN_CHECK(blocks < (int64_t(1) << 31), "frac_f32: too many queries (%lld)", (long long)n);
    if (plan.chain)
    {
        dbuf p32_own;
        float4* p32 = nullptr;
        size_t const np32 = size_t(std::max<int64_t>(plan.npts, 1)) * 2;
        if (ws)
            AGWN_TRY(ws->take_n(np32, &p32));
        else
        {
            AGWN_TRY(p32_own.alloc(np32 * sizeof(float4), stream));
            p32 = p32_own.as<float4>();
        }
        if (plan.npts > 0)
        {
            k_prep_stream<<<dim3(unsigned((plan.npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(
                plan.pts(), plan.npts, R, nullptr, p32);
            AGWN_HIP(hipGetLastError());
        }
        if (split_small && blocks < kSplitBelow && plan.walk_starts.size() > 1)
        {
            // split the stream at walk starts into about kSplitTiles / blocks chunks
            int64_t const want = std::min<int64_t>({int64_t(32), (kSplitTiles + blocks - 1) / blocks,
                                                    int64_t(plan.walk_starts.size())});
            // the cuts are staged in host memory that outlives the copy: the workspace's (valid until its next reset,
            // which follows the owner's synchronisation), else this call's, waited for below
            std::vector<int64_t> own_cuts;
            std::vector<int64_t>& cuts = ws ? ws->host : own_cuts;
            cuts.assign(1, 0);
            for (int64_t c = 1; c < want; ++c)
            {
                auto it = std::lower_bound(plan.walk_starts.begin(), plan.walk_starts.end(), c * plan.npts / want);
                if (it != plan.walk_starts.end() && *it > cuts.back())
                    cuts.push_back(*it);
            }
            cuts.push_back(plan.npts);
            int const chunks = int(cuts.size() - 1);
            dbuf dc_own, part_own;
            int64_t* dc = nullptr;
            double* part = nullptr;
            if (ws)
            {
                AGWN_TRY(ws->take_n(cuts.size(), &dc));
                AGWN_TRY(ws->take_n(size_t(chunks) * size_t(n), &part));
            }
            else
            {
                AGWN_TRY(dc_own.alloc(cuts.size() * sizeof(int64_t), stream));
                AGWN_TRY(part_own.alloc(size_t(chunks) * size_t(n) * sizeof(double), stream));
                dc = dc_own.as<int64_t>();
                part = part_own.as<double>();
            }
            AGWN_HIP(hipMemcpyAsync(dc, cuts.data(), cuts.size() * sizeof(int64_t), hipMemcpyHostToDevice, stream));
            k_frac_f32_chain_part<kQ, Src><<<dim3(unsigned(blocks), unsigned(chunks)), dim3(kBlock), 0, stream>>>(
                pos, n, p32, dc, R, P, kTau, plan.far, part);
            AGWN_HIP(hipGetLastError());
            k_chain_sum<<<dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(part, chunks, n,
                                                                                                counts, out);
            AGWN_HIP(hipGetLastError());
            if (!ws)
                AGWN_HIP(hipStreamSynchronize(stream)); // `cuts` is this call's host vector
            return ok;
        }
        k_frac_f32_chain<kQ, Src><<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(
            pos, n, p32, plan.npts, R, P, kTau, plan.far, counts, out);
        AGWN_HIP(hipGetLastError());
        return ok;
    }
    int64_t const npad = ((nseg + kChunk - 1) / kChunk) * kChunk;
    dbuf s32, s64;
    AGWN_TRY(s32.alloc(size_t(std::max<int64_t>(npad, kChunk)) * sizeof(seg_f32), stream));
    AGWN_TRY(s64.alloc(size_t(std::max<int64_t>(npad, kChunk)) * sizeof(seg_f64), stream));
    if (npad > 0)
    {
        k_prep_segments<float><<<dim3(unsigned((npad + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(
            seg7, nseg, npad, R, s32.as<seg_f32>(), s64.as<seg_f64>());
        AGWN_HIP(hipGetLastError());
    }
    k_frac_f32<kQ, Src><<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(pos, n, s32.as<seg_f32>(),
                                                                           s64.as<seg_f64>(), npad, R, P, kTau,
                                                                           plan.far, counts, out);
    AGWN_HIP(hipGetLastError());
    return ok;
}

// The boundary arrives on the device: a host copy is planned per call (about a millisecond for 30k segments).
template <class Src>
int frac_f32_src(Src const& pos, int64_t n, float const* seg7, int64_t nseg, double const x0[3],
                 int32_t const* counts, float* out, hipStream_t stream)
{
    if (n == 0)
        return ok;
    std::vector<float> h(static_cast<size_t>(nseg) * 7);
    if (nseg > 0)
    {
        AGWN_HIP(hipMemcpyAsync(h.data(), seg7, h.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
        AGWN_HIP(hipStreamSynchronize(stream));
    }
    frac_plan plan;
    AGWN_TRY(frac_plan_build(h.data(), nseg, plan, stream));
    return frac_f32_planned(pos, n, plan, seg7, nseg, x0, counts, out, stream);
}

template <class Src>
int frac_f64_ws(Src const& pos, int64_t n, double const* seg7, int64_t nseg, double const x0[3],
                int32_t const* counts, double* out, seg_f64* ws, hipStream_t stream)
{
    if (n == 0)
        return ok;
    int64_t const blocks = (n + kBlock - 1) / kBlock;
    AGWN_CHECK(blocks < (int64_t(1) << 31), "frac_f64: too many queries (%lld)", (long long)n);
    frame const R = make_frame_impl(x0[0], x0[1], x0[2]);
    if (nseg > 0)
    {
        k_prep_segments<double><<<dim3(unsigned((nseg + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(
            seg7, nseg, nseg, R, nullptr, ws);
        AGWN_HIP(hipGetLastError());
    }
    pole const P = make_pole(x0[0], x0[1], x0[2]);
    k_frac_f64<Src><<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(pos, n, ws, nseg, R, P, counts, out);
    AGWN_HIP(hipGetLastError());
    return ok;
}

template <class Src>
int frac_f64_src(Src const& pos, int64_t n, double const* seg7, int64_t nseg, double const x0[3],
                 int32_t const* counts, double* out, hipStream_t stream, bool chain)
{
    if (n == 0)
        return ok;
    if (chain && nseg > 0)
    {
        // the boundary as a double chain stream when its weights are small integers (planned on the host per call)
        std::vector<double> h(size_t(nseg) * 7);
        AGWN_HIP(hipMemcpyAsync(h.data(), seg7, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        AGWN_HIP(hipStreamSynchronize(stream));
        std::vector<pt64> sp;
        sp.reserve(size_t(nseg + nseg / 8 + 16));
        int64_t nunit = 0;
        if (chain_walk(h.data(), nseg, nunit,
                       [&](double const* p, int flag) { sp.push_back(pt64{p[0], p[1], p[2], double(flag)}); }))
        {
            int64_t const npts = int64_t(sp.size());
            int64_t const blocks = (n + kBlock - 1) / kBlock;
            AGWN_CHECK(blocks < (int64_t(1) << 31), "frac_f64: too many queries (%lld)", (long long)n);
            dbuf raw, rot;
            AGWN_TRY(raw.alloc(size_t(std::max<int64_t>(npts, 1)) * sizeof(pt64), stream));
            AGWN_TRY(rot.alloc(size_t(std::max<int64_t>(npts, 1)) * sizeof(pt64), stream));
            AGWN_HIP(hipMemcpyAsync(raw.p, sp.data(), sp.size() * sizeof(pt64), hipMemcpyHostToDevice, stream));
            frame const R = make_frame_impl(x0[0], x0[1], x0[2]);
            pole const P = make_pole(x0[0], x0[1], x0[2]);
            if (npts > 0)
            {
                k_prep_stream64<<<dim3(unsigned((npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(
                    raw.as<pt64>(), npts, R, rot.as<pt64>());
                AGWN_HIP(hipGetLastError());
            }
            k_frac_f64_chain<Src><<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(pos, n, rot.as<pt64>(), npts,
                                                                                      R, P, counts, out);
            AGWN_HIP(hipGetLastError());
            AGWN_HIP(hipStreamSynchronize(stream)); // `sp` is a host vector
            return ok;
        }
    }
    dbuf s64;
    AGWN_TRY(s64.alloc(size_t(std::max<int64_t>(nseg, 1)) * sizeof(seg_f64), stream));
    return frac_f64_ws(pos, n, seg7, nseg, x0, counts, out, s64.as<seg_f64>(), stream);
}

// (queries per lane, block, LDS stage in blocks of points) shapes of the batched kernel offered to the autotuner
constexpr int kBatchShapes[][3] = {{1, 256, 1}, {2, 256, 1}, {4, 256, 1}, {2, 128, 1}, {4, 128, 1},
                                   {1, 256, 2}, {2, 256, 2}, {4, 256, 2}, {2, 128, 2}, {4, 128, 2},
                                   {1, 256, 4}, {2, 256, 4}, {4, 256, 4}, {2, 128, 4}, {4, 128, 4}};
constexpr int kNumBatchShapes = int(sizeof(kBatchShapes) / sizeof(kBatchShapes[0]));
static_assert(kNumBatchShapes == AGWN_BATCH_FRAC_SHAPES && kNumBatchShapes <= 32, "the header's count of shapes");

int frac_batch_variants() { return kNumBatchShapes; }

void frac_batch_shape(int v, int& q, int& block, int& stage)
{
    q = kBatchShapes[v][0];
    block = kBatchShapes[v][1];
    stage = kBatchShapes[v][2];
}

int frac_batch_blocks_per_cu(int v)
{
    int const b = kBatchShapes[v][1], lds = 32 * kBatchShapes[v][2] * b + 64; // tile + the work counter
    int const waves = b == 256 ? 8 : 16; // 32 waves per CU at most
    return std::max(1, std::min(waves, (160 * 1024) / lds));
}

int frac_stream_rotate(float4 const* in, int64_t n, double const x0[3], int16_t const* pexp, float4* out,
                       hipStream_t s)
{
    if (n == 0)
        return ok;
    frame const R = make_frame_impl(x0[0], x0[1], x0[2]);
    k_prep_stream<<<dim3(unsigned((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s>>>(in, n, R, pexp, out);
    AGWN_HIP(hipGetLastError());
    return ok;
}

namespace
{
struct batch_args
{
    float const* pos;
    batch_frac_desc const* meshes;
    int32_t const* item_mesh;
    int64_t const* item_q0;
    int64_t nitems;
    unsigned long long* counter;
    float4 const* pts;
    frame R;
    pole P;
    int32_t const* counts;
    float* out;
};

template <int V>
void batch_launch_shape(dim3 g, hipStream_t s, batch_args const& a)
{
    constexpr int Q = kBatchShapes[V][0], B = kBatchShapes[V][1], S = kBatchShapes[V][2];
    k_frac_f32_chain_batch<Q, B, S><<<g, dim3(B), 0, s>>>(a.pos, a.meshes, a.item_mesh, a.item_q0, a.nitems,
                                                          a.counter, a.pts, a.R, a.P, kTau, a.counts, a.out);
}

template <int... V>
void batch_dispatch(std::integer_sequence<int, V...>, int v, dim3 g, hipStream_t s, batch_args const& a)
{
    ((v == V ? batch_launch_shape<V>(g, s, a) : void()), ...);
}
} // namespace

int frac_batch_launch(int v, int blocks, float const* pos, batch_frac_desc const* meshes, int32_t const* item_mesh,
                      int64_t const* item_q0, int64_t nitems, unsigned long long* counter, float4 const* pts,
                      double const x0[3], int32_t const* counts, float* out, hipStream_t s)
{
    AGWN_CHECK(v >= 0 && v < frac_batch_variants(), "unknown batch kernel shape %d", v);
    batch_args const a{pos, meshes, item_mesh, item_q0, nitems, counter, pts, make_frame_impl(x0[0], x0[1], x0[2]),
                       make_pole(x0[0], x0[1], x0[2]), counts, out};
    batch_dispatch(std::make_integer_sequence<int, kNumBatchShapes>{}, v, dim3(static_cast<unsigned>(blocks)), s, a);
    AGWN_HIP(hipGetLastError());
    return ok;
}

template int frac_f32_src<src_array<float>>(src_array<float> const&, int64_t, float const*, int64_t, double const*,
                                            int32_t const*, float*, hipStream_t);
template int frac_f32_planned<src_array<float>>(src_array<float> const&, int64_t, frac_plan const&, float const*,
                                                int64_t, double const*, int32_t const*, float*, hipStream_t, bool,
                                                arena*);
template int frac_f32_planned<src_grid>(src_grid const&, int64_t, frac_plan const&, float const*, int64_t,
                                        double const*, int32_t const*, float*, hipStream_t, bool, arena*);
template int frac_f32_src<src_grid>(src_grid const&, int64_t, float const*, int64_t, double const*, int32_t const*,
                                    float*, hipStream_t);
template int frac_f64_src<src_array<double>>(src_array<double> const&, int64_t, double const*, int64_t,
                                             double const*, int32_t const*, double*, hipStream_t, bool);
template int frac_f64_ws<src_array<double>>(src_array<double> const&, int64_t, double const*, int64_t, double const*,
                                            int32_t const*, double*, seg_f64*, hipStream_t);
} // namespace agwn
