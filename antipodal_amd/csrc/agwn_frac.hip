// Fractional GWN term on gfx950: the boundary sum of gwn_mesh.hh:34-46 / math/common.hh:181-189, batched.
//
// Work decomposition. One lane owns Q query points; a 256-lane workgroup owns 256*Q consecutive queries and walks the
// whole boundary. The boundary is pre-rotated once per call into the pole frame, packed 32 B per segment, and read
// with wave-uniform addresses, so each segment arrives in SGPRs through one scalar load and feeds the VALU of all 64
// lanes: no LDS traffic and no bank conflicts in the inner loop.
//
// Pole frame: rows e1 = x1/|x1| (x1 = -x0), e2, e3 orthonormal and right-handed; coordinates (X, y, z) of
// v = endpoint - query. With s = |x1| -- NOT assumed to be 1: a normalised float x0 is off unit length by ~1e-8, and
// the reference's formula is not invariant to |x1| (near the pole the difference reaches 1e-4 rad per pair), so the
// reference's function is evaluated exactly as written:
//   num = x1 . (v0 x v1)                             = s (y0 z1 - z0 y1)
//   den = l0 l1 + l1 x1.v0 + l0 x1.v1 + v0.v1        = l0 l1 + s (l1 X0 + l0 X1) + X0 X1 + y0 y1 + z0 z1
//   area += w * atan2(num, den)
// FP32 kernel: the pair is evaluated in FP32 and accumulated in FP64 (one convert + one FP64 FMA per pair; cheaper
// than FP32 two-sum compensation and strictly more accurate). A pair whose (num, den) is small against l0*l1 is
// ill-conditioned -- the query is near the segment, or an endpoint is seen almost along the pole -- and is
// re-evaluated in FP64 from the FP64 copy of the segment, with the cancellation-free
//   l + s X = (X >= 0 ? l + X : (y^2 + z^2) / (l - X)) + (s - 1) X,
//   den = (l0 + s X0)(l1 + s X1) + (1 - s^2) X0 X1 + y0 y1 + z0 z1,
// then scaled into FP32 range and finished with the FP32 atan2 (the accurate num and den are what matters).
// FP64 kernel: every pair in FP64 with that cancellation-free form and two-sum compensated accumulation.
//
// FP32 chain kernel (boundaries whose weights are small integers -- build_boundary_segments' are +-1): the boundary
// is a stream of points in which consecutive unit segments share endpoints, so each query computes v = P - q and
// l = |v| once per point rather than twice per segment; and the atan2 of two consecutive well-conditioned pairs is
// taken once, of the product of their complex numbers (den + i num): with den > 0 each angle lies in (-pi/2, pi/2),
// so the product's argument is the exact sum of the two with no 2 pi wrap. Pairs with den <= 0 or ill-conditioned
// ones take their own atan2 (FP32) or the FP64 path, as above, and enter the product as 1.

#include "agwn_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

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
            s32[i] = seg_f32{1.0e6f, 0.0f, 0.0f, 1.0e6f, 0.0f, 0.0f, 0.0f, 0.0f};
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
        s32[i] = seg_f32{float(rax), float(ray), float(raz), float(rbx), float(rby), float(rbz), float(w), 0.0f};
    s64[i] = seg_f64{rax, ray, raz, rbx, rby, rbz, w, 0.0};
}

// |x1| and the exact small quantities derived from it, for the FP64 evaluation
struct pole
{
    double s;   // |x1|
    double sm1; // s - 1
    double c;   // 1 - s^2
};

// One pair in FP64, cancellation-free denominator. v0 = a - p, v1 = b - p in the pole frame.
__device__ __forceinline__ double pair_f64(double v0x, double v0y, double v0z, double v1x, double v1y, double v1z,
                                           pole const& P)
{
    double const yz0 = v0y * v0y + v0z * v0z;
    double const yz1 = v1y * v1y + v1z * v1z;
    double const l0 = sqrt(v0x * v0x + yz0);
    double const l1 = sqrt(v1x * v1x + yz1);
    double const lp0 = (v0x >= 0.0 ? l0 + v0x : yz0 / (l0 - v0x)) + P.sm1 * v0x;
    double const lp1 = (v1x >= 0.0 ? l1 + v1x : yz1 / (l1 - v1x)) + P.sm1 * v1x;
    double const num = P.s * (v0y * v1z - v0z * v1y);
    double const den = lp0 * lp1 + (P.c * v0x * v1x + (v0y * v1y + v0z * v1z));
    return atan2(num, den);
}

// atan2 for the FP32 kernel: octant reduction to r = min(|y|,|x|) / max(|y|,|x|) in [0, 1] and a degree-15 odd
// minimax polynomial atan(r) ~ r P(r^2) fitted here (Lawson's iteratively reweighted least squares on [0, 1]),
// |error| <= 1.1e-7 evaluated in FP32. Quadrant fix-ups follow C's atan2 (signed zeros included); no NaN/inf inputs
// reach it (num and den are finite sums of finite products).
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

constexpr int kChunk = 4;          // the boundary is padded to a multiple of this (zero-weight segments)
constexpr int kTile = kBlock;      // segments staged in LDS per pass: one 32-byte segment per lane

template <int Q, class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f32(Src const src, int64_t n, seg_f32 const* __restrict__ seg, seg_f64 const* __restrict__ seg64,
               int64_t nseg_padded, frame R, pole P, float tau, int32_t const* __restrict__ counts,
               float* __restrict__ out)
{
    __shared__ seg_f32 tile[kTile];
    int64_t const base = int64_t(blockIdx.x) * (kBlock * Q) + threadIdx.x;
    float px[Q], py[Q], pz[Q];
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        double x, y, z;
        query_f64(src, base + int64_t(q) * kBlock, n, R, x, y, z);
        px[q] = float(x);
        py[q] = float(y);
        pz[q] = float(z);
        acc[q] = 0.0;
    }
    float const sf = float(P.s);

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
                float const v0x = g.ax - px[q], v0y = g.ay - py[q], v0z = g.az - pz[q];
                float const v1x = g.bx - px[q], v1y = g.by - py[q], v1z = g.bz - pz[q];
                float const l0 = __builtin_amdgcn_sqrtf(v0x * v0x + v0y * v0y + v0z * v0z);
                float const l1 = __builtin_amdgcn_sqrtf(v1x * v1x + v1y * v1y + v1z * v1z);
                float const num = sf * (v0y * v1z - v0z * v1y);
                float const den = l0 * l1 + sf * (l1 * v0x + l0 * v1x) + (v0x * v1x + v0y * v1y + v0z * v1z);
                double a;
                if (__builtin_expect(__builtin_fabsf(num) + __builtin_fabsf(den) < tau * (l0 * l1), 0))
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

// Rotates the point stream into the pole frame: FP32 (flag kept in .w) and FP64 copies.
__global__ void __launch_bounds__(kBlock) k_prep_stream(float4 const* __restrict__ in, int64_t n, frame R,
                                                       float4* __restrict__ out32, double4* __restrict__ out64)
{
    int64_t const i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    float4 const g = in[i];
    double const x = double(g.x), y = double(g.y), z = double(g.z);
    double const rx = R.r[0] * x + R.r[1] * y + R.r[2] * z;
    double const ry = R.r[3] * x + R.r[4] * y + R.r[5] * z;
    double const rz = R.r[6] * x + R.r[7] * y + R.r[8] * z;
    out32[i] = make_float4(float(rx), float(ry), float(rz), g.w);
    out64[i] = make_double4(rx, ry, rz, 0.0);
}

template <int Q, class Src>
__global__ void __launch_bounds__(kBlock)
    k_frac_f32_chain(Src const src, int64_t n, float4 const* __restrict__ pts, double4 const* __restrict__ pts64,
                     int64_t npts, frame R, pole P, float tau, int32_t const* __restrict__ counts,
                     float* __restrict__ out)
{
    __shared__ float4 tile[kTile];
    int64_t const base = int64_t(blockIdx.x) * (kBlock * Q) + threadIdx.x;
    float px[Q], py[Q], pz[Q];
    float vx[Q], vy[Q], vz[Q], vl[Q]; // v and |v| of the previous stream point
    float pn[Q], pd[Q];               // the pending pair's (num, den), when `pending`
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        double x, y, z;
        query_f64(src, base + int64_t(q) * kBlock, n, R, x, y, z);
        px[q] = float(x);
        py[q] = float(y);
        pz[q] = float(z);
        acc[q] = 0.0;
        vx[q] = vy[q] = vz[q] = vl[q] = 0.0f;
        pn[q] = 0.0f;
        pd[q] = 1.0f;
    }
    float const sf = float(P.s);
    bool pending = false; // uniform: the segment structure is the same for every lane

    for (int64_t t0 = 0; t0 < npts; t0 += kTile)
    {
        __syncthreads(); // the previous tile has been consumed by every wave
        if (t0 + threadIdx.x < npts)
            tile[threadIdx.x] = pts[t0 + threadIdx.x];
        __syncthreads();
        int const cnt = int(npts - t0 < kTile ? npts - t0 : kTile);
        for (int j = 0; j < cnt; ++j)
        {
            float4 const g = tile[j]; // same address in every lane: an LDS broadcast
            bool const seg = __builtin_amdgcn_readfirstlane(__float_as_int(g.w)) != 0;
#pragma unroll
            for (int q = 0; q < Q; ++q)
            {
                float const nx = g.x - px[q], ny = g.y - py[q], nz = g.z - pz[q];
                float const nl = __builtin_amdgcn_sqrtf(nx * nx + ny * ny + nz * nz);
                if (seg)
                {
                    float num = sf * (vy[q] * nz - vz[q] * ny);
                    float den = vl[q] * nl + sf * (nl * vx[q] + vl[q] * nx) + (vx[q] * nx + vy[q] * ny + vz[q] * nz);
                    bool const ill = __builtin_fabsf(num) + __builtin_fabsf(den) < tau * (vl[q] * nl);
                    if (__builtin_expect(ill || !(den > 0.0f), 0))
                    {
                        double a;
                        if (ill)
                        {
                            double4 const h0 = pts64[t0 + j - 1], h1 = pts64[t0 + j];
                            double dx, dy, dz;
                            query_f64(src, base + int64_t(q) * kBlock, n, R, dx, dy, dz);
                            a = double(pair_f64_atan_f32(h0.x - dx, h0.y - dy, h0.z - dz, h1.x - dx, h1.y - dy,
                                                         h1.z - dz, P));
                        }
                        else
                        {
                            a = double(atan2_f32(num, den));
                        }
                        acc[q] += a;
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
                vl[q] = nl;
            }
            if (seg)
                pending = !pending;
        }
    }

#pragma unroll
    for (int q = 0; q < Q; ++q)
    {
        if (pending)
            acc[q] += double(atan2_f32(pn[q], pd[q]));
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
#pragma unroll 2
    for (int64_t s = 0; s < nseg; ++s)
    {
        seg_f64 const g = seg[s];
        double const a = pair_f64(g.ax - px, g.ay - py, g.az - pz, g.bx - px, g.by - py, g.bz - pz, P);
        double t, e;
        two_sum(sum, g.w * a, t, e);
        sum = t;
        comp += e;
    }
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
constexpr int kQ = 2;
} // namespace

frame make_pole_frame(double x0x, double x0y, double x0z) { return make_frame_impl(x0x, x0y, x0z); }

// Unit-weight expansion and chain ordering of a host boundary into a point stream (x, y, z, flag). Weights must be
// integers of magnitude <= 16 (0 drops the segment); otherwise returns false. Greedy walks in input order: start at
// the first unused unit segment, follow any unused segment leaving the current end point (exact coordinate match),
// stop when there is none. Every unit segment appears exactly once, as a flag-1 point after its start point.
bool chain_stream(float const* h, int64_t nseg, std::vector<float4>& stream_pts, int64_t& nunit)
{
    stream_pts.clear();
    nunit = 0;
    int64_t total = 0;
    for (int64_t i = 0; i < nseg; ++i)
    {
        float const w = h[7 * i + 6];
        if (!(w == std::rint(w)) || std::fabs(w) > 16.0f)
            return false; // general weights: the per-segment kernel
        total += int64_t(std::fabs(w));
    }
    struct key
    {
        uint32_t b[3];
        bool operator<(key const& o) const
        {
            return b[0] != o.b[0] ? b[0] < o.b[0] : (b[1] != o.b[1] ? b[1] < o.b[1] : b[2] < o.b[2]);
        }
        bool operator==(key const& o) const { return b[0] == o.b[0] && b[1] == o.b[1] && b[2] == o.b[2]; }
    };
    auto key_of = [](float const* p) {
        key k;
        std::memcpy(k.b, p, sizeof(k.b));
        return k;
    };
    // unit edges (from, to): pointers to the endpoint coordinates
    std::vector<float const*> from, to;
    from.reserve(size_t(total));
    to.reserve(size_t(total));
    for (int64_t i = 0; i < nseg; ++i)
    {
        float const* g = h + 7 * i;
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
    auto next_from = [&](float const* p) -> int64_t {
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
    stream_pts.reserve(size_t(E + E / 8 + 16));
    for (int64_t e0 = 0; e0 < E; ++e0)
    {
        if (used[size_t(e0)])
            continue;
        float const* a = from[size_t(e0)];
        stream_pts.push_back(make_float4(a[0], a[1], a[2], 0.0f));
        int64_t e = e0;
        while (e >= 0)
        {
            used[size_t(e)] = 1;
            float const* b = to[size_t(e)];
            stream_pts.push_back(make_float4(b[0], b[1], b[2], 1.0f));
            e = next_from(b);
        }
    }
    nunit = E;
    return true;
}

int frac_plan_build(float const* h, int64_t nseg, frac_plan& plan, hipStream_t stream)
{
    plan.chain = false;
    plan.npts = plan.nunit = 0;
    std::vector<float4> stream_pts;
    int64_t nunit = 0;
    if (!chain_stream(h, nseg, stream_pts, nunit))
        return ok;
    plan.npts = int64_t(stream_pts.size());
    plan.nunit = nunit;
    AGWN_TRY(plan.buf.alloc(std::max<size_t>(stream_pts.size(), 1) * sizeof(float4), stream));
    if (!stream_pts.empty())
        AGWN_HIP(hipMemcpyAsync(plan.buf.p, stream_pts.data(), stream_pts.size() * sizeof(float4),
                                hipMemcpyHostToDevice, stream));
    AGWN_HIP(hipStreamSynchronize(stream)); // the host vector goes out of scope
    plan.chain = true;
    return ok;
}

template <class Src>
int frac_f32_planned(Src const& pos, int64_t n, frac_plan const& plan, float const* seg7, int64_t nseg,
                     double const x0[3], int32_t const* counts, float* out, hipStream_t stream)
{
    if (n == 0)
        return ok;
    frame const R = make_frame_impl(x0[0], x0[1], x0[2]);
    pole const P = make_pole(x0[0], x0[1], x0[2]);
    int64_t const per_block = int64_t(kBlock) * kQ;
    int64_t const blocks = (n + per_block - 1) / per_block;
    AGWN_CHECK(blocks < (int64_t(1) << 31), "frac_f32: too many queries (%lld)", (long long)n);
    if (plan.chain)
    {
        dbuf p32, p64;
        AGWN_TRY(p32.alloc(size_t(std::max<int64_t>(plan.npts, 1)) * sizeof(float4), stream));
        AGWN_TRY(p64.alloc(size_t(std::max<int64_t>(plan.npts, 1)) * sizeof(double4), stream));
        if (plan.npts > 0)
        {
            k_prep_stream<<<dim3(unsigned((plan.npts + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(
                plan.pts(), plan.npts, R, p32.as<float4>(), p64.as<double4>());
            AGWN_HIP(hipGetLastError());
        }
        k_frac_f32_chain<kQ, Src><<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(
            pos, n, p32.as<float4>(), p64.as<double4>(), plan.npts, R, P, kTau, counts, out);
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
                                                                           counts, out);
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
int frac_f64_src(Src const& pos, int64_t n, double const* seg7, int64_t nseg, double const x0[3],
                 int32_t const* counts, double* out, hipStream_t stream)
{
    if (n == 0)
        return ok;
    frame const R = make_frame_impl(x0[0], x0[1], x0[2]);
    dbuf s64;
    AGWN_TRY(s64.alloc(size_t(std::max<int64_t>(nseg, 1)) * sizeof(seg_f64), stream));
    if (nseg > 0)
    {
        k_prep_segments<double><<<dim3(unsigned((nseg + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream>>>(
            seg7, nseg, nseg, R, nullptr, s64.as<seg_f64>());
        AGWN_HIP(hipGetLastError());
    }
    int64_t const blocks = (n + kBlock - 1) / kBlock;
    AGWN_CHECK(blocks < (int64_t(1) << 31), "frac_f64: too many queries (%lld)", (long long)n);
    pole const P = make_pole(x0[0], x0[1], x0[2]);
    k_frac_f64<Src><<<dim3(unsigned(blocks)), dim3(kBlock), 0, stream>>>(pos, n, s64.as<seg_f64>(), nseg, R, P,
                                                                         counts, out);
    AGWN_HIP(hipGetLastError());
    return ok;
}

template int frac_f32_src<src_array<float>>(src_array<float> const&, int64_t, float const*, int64_t, double const*,
                                            int32_t const*, float*, hipStream_t);
template int frac_f32_planned<src_array<float>>(src_array<float> const&, int64_t, frac_plan const&, float const*,
                                                int64_t, double const*, int32_t const*, float*, hipStream_t);
template int frac_f32_planned<src_grid>(src_grid const&, int64_t, frac_plan const&, float const*, int64_t,
                                        double const*, int32_t const*, float*, hipStream_t);
template int frac_f32_src<src_grid>(src_grid const&, int64_t, float const*, int64_t, double const*, int32_t const*,
                                    float*, hipStream_t);
template int frac_f64_src<src_array<double>>(src_array<double> const&, int64_t, double const*, int64_t,
                                             double const*, int32_t const*, double*, hipStream_t);
} // namespace agwn
