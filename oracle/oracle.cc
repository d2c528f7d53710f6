// =====================================================================================================================
// ORACLE -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of antipodal's batch mesh-GWN path, written from the reference's algorithm (file:line cited on each
// function; paths are relative to the reference repository root). It is the checker for the HIP product in
// antipodal_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load liboracle.so, and only
// to check or to time a CPU baseline. The product never links, loads or calls anything in this directory.
//
// Pinning (what checks each function, so that a plausible mistake fails a test):
//   pair_half_area      closed forms of test/test_common.cc (octant pi/4, equator loop pi, flips, scale invariance,
//                       degenerate zeros) + golden vectors from the reference's own code (oracle/_ref).
//   frac_single         golden vectors from the reference's eval_gwnr_mesh_batch_fractional (bit-level for f64);
//                       frac + int == per-triangle solid-angle sum on open meshes.
//   naive_count         golden vectors from the reference's NaiveIntersector (bit-exact); closed cube counts 1/0.
//   gwn_baseline        golden vectors from the reference's test/gwn_baseline.hh; closed mesh gives {0,1}.
//   boundary            multiset equality with the reference's build_boundary_segments (golden).
//   edge_det/edge_sign/atan2_num/project_into_tri
//                       the reference's own robust_predicates.hh compiled in oracle/_ref (random + lattice cases);
//                       int32 vs int64 vs double agreement on 2^-20 lattices (the reference's own test scheme).
//   robust_eval         the reference's test_gwn_mesh_robust.cc geometries and thresholds against gwn_baseline;
//                       exact_float fractional == plain f64 fractional with x0 = -x off the grazing set.
//   Exact sign of a double determinant: computed here with a multi-limb integer (independent of the reference's
//   floating-point expansion and of the GPU kernels' expansion), so it pins both.
//
// Build: oracle/Makefile (g++ -std=c++20 -O2 -ffp-contract=off -fopenmp). No -ffast-math: the restatement relies on
// IEEE round-to-nearest, exactly as the reference does.
// =====================================================================================================================

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numbers>
#include <vector>

#include <omp.h>

namespace orc
{
// ---------------------------------------------------------------------------------------------------------------------
// small vector helpers -- the evaluation order of every expression matches the reference's math/common.hh:74-101
// (dot = (ax*bx + ay*by) + az*bz, cross component order), so float results are bit-identical to the reference's.
// ---------------------------------------------------------------------------------------------------------------------
template <class T>
struct V3
{
    T x, y, z;
};

template <class T>
static inline V3<T> load3(T const* p)
{
    return {p[0], p[1], p[2]};
}
template <class T>
static inline V3<T> sub(V3<T> a, V3<T> b)
{
    return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
static inline T dot3(V3<T> a, V3<T> b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
static inline V3<T> cross3(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
static inline T len3(V3<T> a)
{
    return std::sqrt(dot3(a, a));
}

template <class T>
static constexpr T kPi = std::numbers::pi_v<T>;

// math/common.hh:181-189 -- half the signed solid angle of the spherical triangle (x1, v0, v1), unnormalised VOS.
template <class T>
static inline T pair_half_area(V3<T> x1, V3<T> v0, V3<T> v1)
{
    T const l0 = len3(v0);
    T const l1 = len3(v1);
    T const num = dot3(x1, cross3(v0, v1));
    T const den = l0 * l1 + l1 * dot3(x1, v0) + l0 * dot3(x1, v1) + dot3(v0, v1);
    return std::atan2(num, den);
}

// gwn_mesh.hh:34-46 -- sequential weighted sum over the boundary in the given order, divided by 2*pi.
// seg7 = {pos0.xyz, pos1.xyz, weight} per segment (the in-memory layout of weighted_segment3, common.hh:60-65).
template <class T>
static inline T frac_single(V3<T> p, V3<T> x0, T const* seg7, int64_t nseg)
{
    V3<T> const x1{-x0.x, -x0.y, -x0.z};
    T area = T(0);
    for (int64_t s = 0; s < nseg; ++s)
    {
        T const* g = seg7 + 7 * s;
        V3<T> const v0 = sub(load3(g), p);
        V3<T> const v1 = sub(load3(g + 3), p);
        area += g[6] * pair_half_area(x1, v0, v1);
    }
    return area / (T(2) * kPi<T>);
}

// intersector.hh:61-100 -- Moeller-Trumbore over every triangle, same operations in the same order.
template <class T>
static inline int naive_count(V3<T> p, V3<T> dir, T const* verts, int32_t const* idx, int64_t ntri)
{
    int sum = 0;
    for (int64_t t = 0; t < ntri; ++t)
    {
        V3<T> const a = load3(verts + 3 * int64_t(idx[3 * t + 0]));
        V3<T> const b = load3(verts + 3 * int64_t(idx[3 * t + 1]));
        V3<T> const c = load3(verts + 3 * int64_t(idx[3 * t + 2]));
        V3<T> const e1 = sub(b, a);
        V3<T> const e2 = sub(c, a);
        V3<T> const h = cross3(dir, e2);
        T const det = dot3(e1, h);
        if (det == T(0))
            continue;
        T const inv = T(1) / det;
        V3<T> const s = sub(p, a);
        T const u = inv * dot3(s, h);
        if (u < T(0) || u > T(1))
            continue;
        V3<T> const q = cross3(s, e1);
        T const v = inv * dot3(dir, q);
        if (v < T(0) || u + v > T(1))
            continue;
        T const th = inv * dot3(e2, q);
        if (th < T(0))
            continue;
        sum += dot3(cross3(e1, e2), dir) > T(0) ? +1 : -1;
    }
    return sum;
}

// test/gwn_baseline.hh:20-43 -- per-triangle VOS solid angle of normalised vertex directions, / 4*pi.
// With `compensated` the running sum is Neumaier-compensated (used only for large meshes, where the plain running sum
// of ~1e6 terms would itself lose digits); the per-triangle terms are the reference's.
template <class T>
static inline T gwn_baseline(V3<T> p, T const* verts, int32_t const* idx, int64_t ntri, bool compensated)
{
    T sum = T(0), comp = T(0);
    for (int64_t t = 0; t < ntri; ++t)
    {
        V3<T> a = sub(load3(verts + 3 * int64_t(idx[3 * t + 0])), p);
        V3<T> b = sub(load3(verts + 3 * int64_t(idx[3 * t + 1])), p);
        V3<T> c = sub(load3(verts + 3 * int64_t(idx[3 * t + 2])), p);
        T const la = len3(a), lb = len3(b), lc = len3(c);
        a = {a.x / la, a.y / la, a.z / la};
        b = {b.x / lb, b.y / lb, b.z / lb};
        c = {c.x / lc, c.y / lc, c.z / lc};
        T const num = dot3(a, cross3(b, c));
        T const den = T(1) + dot3(a, b) + dot3(b, c) + dot3(c, a);
        T const term = T(2) * std::atan2(num, den);
        if (!compensated)
        {
            sum += term;
        }
        else
        {
            T const s2 = sum + term;
            comp += (std::abs(sum) >= std::abs(term)) ? (sum - s2) + term : (term - s2) + sum;
            sum = s2;
        }
    }
    return (sum + comp) / (T(4) * kPi<T>);
}

// ---------------------------------------------------------------------------------------------------------------------
// math/common.hh:123-174 -- open boundary as weighted oriented segments.
// Net directed count per undirected edge {lo,hi}: +1 for a lo->hi traversal, -1 for hi->lo; zero-sum edges dropped;
// the segment is emitted lo->hi with weight sum if sum > 0, else hi->lo with weight -sum. The reference emits in
// unordered_map order (unspecified); this restatement emits sorted by (lo,hi). Parity is on the multiset.
// ---------------------------------------------------------------------------------------------------------------------
template <class T>
static int64_t boundary(T const* verts, int32_t const* idx, int64_t ntri, T* out7, int64_t cap)
{
    std::vector<std::pair<uint64_t, int>> half;
    half.reserve(size_t(3 * ntri));
    for (int64_t t = 0; t < ntri; ++t)
        for (int e = 0; e < 3; ++e)
        {
            int32_t const u = idx[3 * t + e], v = idx[3 * t + (e + 1) % 3];
            int32_t const lo = std::min(u, v), hi = std::max(u, v);
            half.push_back({(uint64_t(uint32_t(lo)) << 32) | uint32_t(hi), u < v ? 1 : -1});
        }
    std::sort(half.begin(), half.end(), [](auto const& a, auto const& b) { return a.first < b.first; });
    int64_t n = 0;
    for (size_t i = 0; i < half.size();)
    {
        size_t j = i;
        int net = 0;
        while (j < half.size() && half[j].first == half[i].first)
            net += half[j++].second;
        if (net != 0)
        {
            int32_t const lo = int32_t(half[i].first >> 32), hi = int32_t(half[i].first & 0xffffffffu);
            int32_t const a = net > 0 ? lo : hi, b = net > 0 ? hi : lo;
            if (n < cap)
            {
                T* o = out7 + 7 * n;
                for (int k = 0; k < 3; ++k)
                {
                    o[k] = verts[3 * int64_t(a) + k];
                    o[3 + k] = verts[3 * int64_t(b) + k];
                }
                o[6] = T(net > 0 ? net : -net);
            }
            ++n;
        }
        i = j;
    }
    return n;
}

// ---------------------------------------------------------------------------------------------------------------------
// Robust predicates -- math/robust_predicates.hh:225-411.
// Coordinates C in {int32 (int64_grid), int64 (int128_grid), double (exact_float)}; W = the wide type of edge_det.
// ---------------------------------------------------------------------------------------------------------------------
using i128 = __int128;

template <class C>
struct wide_of;
template <>
struct wide_of<int32_t>
{
    using type = int64_t;
};
template <>
struct wide_of<int64_t>
{
    using type = i128;
};
template <>
struct wide_of<double>
{
    using type = double;
};
template <class C>
using W = typename wide_of<C>::type;

template <class C>
static inline W<C> wmul(C a, C b)
{
    return W<C>(a) * W<C>(b);
}

// Exact sign of sum_i coef_i * a_i * b_i for doubles, via a multi-limb integer: every double is m * 2^e with a 53-bit
// integer m, so each product is an integer of at most 106 bits times a power of two; align all products to the
// smallest exponent and add them in base 2^32 limbs wide enough for the whole exponent range.
static int exact_sign_sum_products(int n, double const* a, double const* b, int const* coef)
{
    struct term
    {
        unsigned __int128 mag;
        int exp;
        int sgn;
    };
    term t[8];
    int nt = 0;
    for (int i = 0; i < n; ++i)
    {
        if (a[i] == 0.0 || b[i] == 0.0)
            continue;
        int ea, eb;
        double const fa = std::frexp(std::abs(a[i]), &ea);
        double const fb = std::frexp(std::abs(b[i]), &eb);
        uint64_t const ma = uint64_t(std::ldexp(fa, 53));
        uint64_t const mb = uint64_t(std::ldexp(fb, 53));
        int const s = ((a[i] < 0) != (b[i] < 0) ? -1 : 1) * coef[i];
        t[nt++] = {(unsigned __int128)ma * mb, ea + eb - 106, s};
    }
    if (nt == 0)
        return 0;
    int emin = t[0].exp, emax = t[0].exp;
    for (int i = 1; i < nt; ++i)
    {
        emin = std::min(emin, t[i].exp);
        emax = std::max(emax, t[i].exp);
    }
    int const nlimb = (emax - emin + 128) / 32 + 4;
    std::vector<int64_t> limb(size_t(nlimb), 0);
    for (int i = 0; i < nt; ++i)
    {
        int const shift = t[i].exp - emin;
        int const base = shift / 32, off = shift % 32;
        // mag << off spans at most 107 + 31 bits: split into 32-bit chunks.
        unsigned __int128 m = t[i].mag;
        uint64_t chunk[6] = {0, 0, 0, 0, 0, 0};
        // shift by `off` across the chunks
        unsigned __int128 lo = m << off; // low 128 bits of the shifted value
        unsigned __int128 hi = off ? (m >> (128 - off)) : 0;
        for (int k = 0; k < 4; ++k)
            chunk[k] = uint64_t((lo >> (32 * k)) & 0xffffffffu);
        chunk[4] = uint64_t(hi & 0xffffffffu);
        chunk[5] = uint64_t((hi >> 32) & 0xffffffffu);
        for (int k = 0; k < 6; ++k)
            limb[size_t(base + k)] += t[i].sgn * int64_t(chunk[k]);
    }
    for (int k = 0; k + 1 < nlimb; ++k)
    {
        int64_t const carry = limb[size_t(k)] >> 32; // arithmetic shift = floor division
        limb[size_t(k)] -= carry * (int64_t(1) << 32);
        limb[size_t(k + 1)] += carry;
    }
    for (int k = nlimb - 1; k >= 0; --k)
        if (limb[size_t(k)] != 0)
            return limb[size_t(k)] > 0 ? 1 : -1;
    return 0;
}

// robust_predicates.hh:272-308 (edge_det): the 2x2 determinant of (q - q0, q1 - q0) in the (y,z) plane.
// Integers: exact. Doubles: exact sign; magnitude = the rounded double estimate (p1 + p2), raised to DBL_MIN when the
// exact value is nonzero and the estimate underflows or carries the wrong sign (the reference's filter returns the
// estimate itself exactly when its sign is certain, and max(|estimate|, DBL_MIN) with the exact sign otherwise).
template <class C>
static inline W<C> edge_det(V3<C> q, V3<C> q0, V3<C> q1)
{
    if constexpr (std::is_same_v<C, double>)
    {
        double const p1 = (q.y - q0.y) * (q0.z - q1.z);
        double const p2 = (q.z - q0.z) * (q1.y - q0.y);
        double const det = p1 + p2;
        // det = qy*q0z - qy*q1z + q0y*q1z + qz*q1y - qz*q0y - q0z*q1y in raw coordinates
        double const a[6] = {q.y, q.y, q0.y, q.z, q.z, q0.z};
        double const b[6] = {q0.z, q1.z, q1.z, q1.y, q0.y, q1.y};
        int const coef[6] = {+1, -1, +1, +1, -1, -1};
        int const s = exact_sign_sum_products(6, a, b, coef);
        if (s == 0)
            return 0.0;
        return std::copysign(std::max(std::abs(det), std::numeric_limits<double>::min()), double(s));
    }
    else
    {
        return wmul<C>(C(q.y - q0.y), C(q0.z - q1.z)) + wmul<C>(C(q.z - q0.z), C(q1.y - q0.y));
    }
}

// robust_predicates.hh:314-331 (edge_sign): sign of edge_det, tie broken by the perturbation eps2 (y) >> eps3 (z).
template <class C>
static inline int edge_sign(V3<C> q, V3<C> q0, V3<C> q1)
{
    W<C> const d = edge_det(q, q0, q1);
    if (d != 0)
        return d > 0 ? 1 : -1;
    if (q0.z != q1.z)
        return q0.z > q1.z ? 1 : -1;
    if (q1.y != q0.y)
        return q1.y > q0.y ? 1 : -1;
    return 0;
}

// robust_predicates.hh:339-356 (atan2_num): num = v0.z*v1.y - v0.y*v1.z with v = q_i - q; exact sign (= -edge_det).
template <class C>
static inline W<C> atan2_num(V3<C> q, V3<C> q0, V3<C> q1)
{
    V3<C> const v0 = sub(q0, q), v1 = sub(q1, q);
    if constexpr (std::is_same_v<C, double>)
    {
        double const t = edge_det(q, q0, q1);
        if (t == 0)
            return 0.0;
        double const num = v0.z * v1.y - v0.y * v1.z;
        return std::copysign(std::max(std::abs(num), std::numeric_limits<double>::min()), -t);
    }
    else
    {
        return wmul<C>(v0.z, v1.y) - wmul<C>(v0.y, v1.z);
    }
}

// robust_predicates.hh:368-375 (dot_yz)
template <class C>
static inline double dot_yz(V3<C> a, V3<C> b)
{
    if constexpr (std::is_same_v<C, double>)
        return a.y * b.y + a.z * b.z;
    else
        return double(wmul<C>(a.y, b.y) + wmul<C>(a.z, b.z));
}

// robust_predicates.hh:382-402 (tri_projection, project_into_tri)
template <class C>
static inline int project_into_tri(V3<C> q, V3<C> q0, V3<C> q1, V3<C> q2, double* w)
{
    int const e01 = edge_sign(q, q0, q1);
    int const e12 = edge_sign(q, q1, q2);
    int const e20 = edge_sign(q, q2, q0);
    w[0] = w[1] = w[2] = 0.0;
    if (e01 == 0 || e01 != e12 || e01 != e20)
        return 0;
    w[0] = double(edge_det(q, q1, q2));
    w[1] = double(edge_det(q, q2, q0));
    w[2] = double(edge_det(q, q0, q1));
    return e01;
}

// ---------------------------------------------------------------------------------------------------------------------
// RobustMeshGwn restatement -- gwn_mesh_robust.hh:170-456, with the Embree culling replaced by a brute-force loop over
// every triangle (Embree only culls conservatively: every triangle the exact predicate accepts is inside its padded
// box, gwn_mesh_robust.hh:82-90, so the brute-force count is the reference's count).
// ---------------------------------------------------------------------------------------------------------------------
template <class C>
struct robust_mesh
{
    double cx = 0, cy = 0, cz = 0; // m_center
    double scale = 1.0;            // m_scale
    int grid_bits = 0;             // 0 = exact_float
    std::vector<V3<C>> tri;        // 3 per triangle, the predicate's copy
    std::vector<V3<C>> seg;        // 2 per segment, stored SWAPPED (pos1, pos0)
    std::vector<double> w;

    // gwn_mesh_robust.hh:334-351 (quantize)
    V3<C> quantize(V3<double> p) const
    {
        if constexpr (std::is_same_v<C, double>)
        {
            return p;
        }
        else
        {
            double const max_q = double(int64_t(1) << (grid_bits + 10));
            double const qx = (p.x - cx) * scale, qy = (p.y - cy) * scale, qz = (p.z - cz) * scale;
            auto const r = [&](double v) { return C(std::llround(std::clamp(v, -max_q, max_q))); };
            return {r(qx), r(qy), r(qz)};
        }
    }
};

// gwn_mesh_robust.hh:177-209 + build_quantization :354-378
template <class C, class T>
static robust_mesh<C> robust_build(int grid_bits, T const* verts, int64_t nv, int32_t const* idx, int64_t ntri,
                                   T const* seg7, int64_t nseg)
{
    robust_mesh<C> m;
    m.grid_bits = grid_bits;
    if (grid_bits > 0 && nv > 0)
    {
        double mn[3] = {double(verts[0]), double(verts[1]), double(verts[2])};
        double mx[3] = {mn[0], mn[1], mn[2]};
        for (int64_t i = 0; i < nv; ++i)
            for (int k = 0; k < 3; ++k)
            {
                mn[k] = std::min(mn[k], double(verts[3 * i + k]));
                mx[k] = std::max(mx[k], double(verts[3 * i + k]));
            }
        double const ex = mx[0] - mn[0], ey = mx[1] - mn[1], ez = mx[2] - mn[2];
        m.cx = mn[0] + ex * 0.5;
        m.cy = mn[1] + ey * 0.5;
        m.cz = mn[2] + ez * 0.5;
        double const half = 0.5 * std::max({ex, ey, ez});
        double const max_coord = double((int64_t(1) << grid_bits) - 1);
        m.scale = half > 0.0 ? max_coord / half : 1.0;
    }
    auto const tod = [&](T const* p) { return V3<double>{double(p[0]), double(p[1]), double(p[2])}; };
    m.tri.reserve(size_t(3 * ntri));
    for (int64_t t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k)
            m.tri.push_back(m.quantize(tod(verts + 3 * int64_t(idx[3 * t + k]))));
    for (int64_t s = 0; s < nseg; ++s)
    {
        m.seg.push_back(m.quantize(tod(seg7 + 7 * s + 3)));
        m.seg.push_back(m.quantize(tod(seg7 + 7 * s + 0)));
        m.w.push_back(double(seg7[7 * s + 6]));
    }
    return m;
}

// gwn_mesh_robust.hh:122-145 + :231-258 -- count of the -x ray, brute force over all triangles.
template <class C>
static inline int robust_count(robust_mesh<C> const& m, V3<C> qi)
{
    int sum = 0;
    size_t const ntri = m.tri.size() / 3;
    for (size_t t = 0; t < ntri; ++t)
    {
        V3<C> const a = m.tri[3 * t], b = m.tri[3 * t + 1], c = m.tri[3 * t + 2];
        double w[3];
        int const s = project_into_tri(qi, a, b, c, w);
        if (s == 0)
            continue;
        double const hx = (w[0] * double(a.x) + w[1] * double(b.x) + w[2] * double(c.x)) / (w[0] + w[1] + w[2]);
        if (double(qi.x) >= hx)
            sum += -s;
    }
    return sum;
}

// gwn_mesh_robust.hh:264-325 -- robust fractional term.
template <class C>
static inline double robust_frac(robust_mesh<C> const& m, V3<C> qi)
{
    constexpr double pi = kPi<double>;
    constexpr double f_norm = 2.0 / (4.0 * pi);
    double half_area = 0.0;
    size_t const nseg = m.w.size();
    for (size_t s = 0; s < nseg; ++s)
    {
        V3<C> const p0 = m.seg[2 * s], p1 = m.seg[2 * s + 1];
        V3<C> const v0 = sub(p0, qi), v1 = sub(p1, qi);
        V3<double> const d0{double(v0.x), double(v0.y), double(v0.z)};
        V3<double> const d1{double(v1.x), double(v1.y), double(v1.z)};
        double const l0 = len3(d0), l1 = len3(d1);
        W<C> const num = atan2_num(qi, p0, p1);
        auto const lpx = [](V3<C> v, double l) { return v.x >= 0 ? l + double(v.x) : dot_yz(v, v) / (l - double(v.x)); };
        double const den = lpx(v0, l0) * lpx(v1, l1) + dot_yz(v0, v1);
        double contrib;
        if (num != 0)
            contrib = std::atan2(double(num), den);
        else if (v0.y == 0 && v0.z == 0 && v0.x < 0 && (v1.y != 0 || v1.z != 0))
            contrib = v1.z != 0 ? std::atan2(double(v1.z), -double(v1.y)) : (v1.y > 0 ? -pi : 0.0);
        else if (v1.y == 0 && v1.z == 0 && v1.x < 0 && (v0.y != 0 || v0.z != 0))
            contrib = v0.z != 0 ? std::atan2(-double(v0.z), -double(v0.y)) : (v0.y > 0 ? pi : 0.0);
        else
            contrib = den < 0.0 ? double(-edge_sign(qi, p0, p1)) * pi : 0.0;
        half_area += m.w[s] * contrib;
    }
    return half_area * f_norm;
}

template <class C, class T>
static void robust_run(int grid_bits, T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* seg7,
                       int64_t nseg, T const* pos, int64_t n, T* out_frac, int32_t* out_count, T* out_eval, int nthreads)
{
    auto const m = robust_build<C, T>(grid_bits, verts, nv, idx, nidx / 3, seg7, nseg);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t i = 0; i < n; ++i)
    {
        V3<C> const qi = m.quantize({double(pos[3 * i]), double(pos[3 * i + 1]), double(pos[3 * i + 2])});
        T const f = T(robust_frac(m, qi));
        int const c = robust_count(m, qi);
        if (out_frac)
            out_frac[i] = f;
        if (out_count)
            out_count[i] = c;
        if (out_eval)
            out_eval[i] = f + T(c);
    }
}
} // namespace orc

using namespace orc;

// =====================================================================================================================
// C ABI for the tests (ctypes). nthreads <= 0 means "all OpenMP threads".
// =====================================================================================================================
static int nt(int n) { return n > 0 ? n : omp_get_max_threads(); }

extern "C" {

float orc_pair_half_area_f32(float const* x1, float const* v0, float const* v1)
{
    return pair_half_area(load3(x1), load3(v0), load3(v1));
}
double orc_pair_half_area_f64(double const* x1, double const* v0, double const* v1)
{
    return pair_half_area(load3(x1), load3(v0), load3(v1));
}

#define ORC_PLAIN(T, SFX)                                                                                               \
    void orc_frac_batch_##SFX(T const* pos, int64_t n, T const* seg7, int64_t nseg, T const* x0, T* out, int nth)     \
    {                                                                                                                  \
        V3<T> const d = load3(x0);                                                                                     \
        _Pragma("omp parallel for schedule(dynamic, 16) num_threads(nt(nth))") for (int64_t i = 0; i < n; ++i)         \
            out[i] = frac_single(load3(pos + 3 * i), d, seg7, nseg);                                                   \
    }                                                                                                                  \
    void orc_naive_count_##SFX(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* pos, int64_t n,  \
                               T const* dir, int32_t* out, int nth)                                                    \
    {                                                                                                                  \
        (void)nv;                                                                                                      \
        V3<T> const d = load3(dir);                                                                                    \
        _Pragma("omp parallel for schedule(dynamic, 16) num_threads(nt(nth))") for (int64_t i = 0; i < n; ++i)         \
            out[i] = naive_count(load3(pos + 3 * i), d, verts, idx, nidx / 3);                                         \
    }                                                                                                                  \
    void orc_full_batch_##SFX(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* pos, int64_t n,   \
                              T const* x0, T const* seg7, int64_t nseg, T* out, int nth)                               \
    {                                                                                                                  \
        (void)nv;                                                                                                      \
        V3<T> const d = load3(x0);                                                                                     \
        _Pragma("omp parallel for schedule(dynamic, 16) num_threads(nt(nth))") for (int64_t i = 0; i < n; ++i)         \
        {                                                                                                              \
            V3<T> const p = load3(pos + 3 * i);                                                                        \
            out[i] = frac_single(p, d, seg7, nseg) + T(naive_count(p, d, verts, idx, nidx / 3));                       \
        }                                                                                                              \
    }                                                                                                                  \
    void orc_gwn_baseline_##SFX(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T const* pos, int64_t n, \
                                T* out, int compensated, int nth)                                                      \
    {                                                                                                                  \
        (void)nv;                                                                                                      \
        _Pragma("omp parallel for schedule(dynamic, 4) num_threads(nt(nth))") for (int64_t i = 0; i < n; ++i)          \
            out[i] = gwn_baseline(load3(pos + 3 * i), verts, idx, nidx / 3, compensated != 0);                         \
    }                                                                                                                  \
    int64_t orc_boundary_##SFX(T const* verts, int64_t nv, int32_t const* idx, int64_t nidx, T* out7, int64_t cap)     \
    {                                                                                                                  \
        (void)nv;                                                                                                      \
        return boundary(verts, idx, nidx / 3, out7, cap);                                                              \
    }

ORC_PLAIN(float, f32)
ORC_PLAIN(double, f64)

// The fractional term of double inputs evaluated in x87 extended precision (64-bit significand): the arbiter for
// FP64 results, whose own rounding near a boundary segment (an ill-conditioned atan2) can reach ~1e-12.
void orc_frac_batch_f64x(double const* pos, int64_t n, double const* seg7, int64_t nseg, double const* x0, double* out,
                         int nth)
{
    using L = long double;
    std::vector<L> s(size_t(7 * nseg));
    for (int64_t i = 0; i < 7 * nseg; ++i)
        s[size_t(i)] = L(seg7[i]);
    V3<L> const d{L(x0[0]), L(x0[1]), L(x0[2])};
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt(nth))
    for (int64_t i = 0; i < n; ++i)
        out[i] = double(frac_single(V3<L>{L(pos[3 * i]), L(pos[3 * i + 1]), L(pos[3 * i + 2])}, d, s.data(), nseg));
}

// --- predicates, for pinning against the reference's robust_predicates.hh ------------------------------------------
#define ORC_PRED(C, SFX)                                                                                               \
    double orc_edge_det_##SFX(C const* q, C const* a, C const* b) { return double(edge_det(load3(q), load3(a), load3(b))); } \
    int orc_edge_det_sign_##SFX(C const* q, C const* a, C const* b)                                                    \
    {                                                                                                                  \
        auto const d = edge_det(load3(q), load3(a), load3(b));                                                         \
        return d > 0 ? 1 : (d < 0 ? -1 : 0);                                                                           \
    }                                                                                                                  \
    int orc_edge_sign_##SFX(C const* q, C const* a, C const* b) { return edge_sign(load3(q), load3(a), load3(b)); }    \
    double orc_atan2_num_##SFX(C const* q, C const* a, C const* b) { return double(atan2_num(load3(q), load3(a), load3(b))); } \
    int orc_atan2_num_sign_##SFX(C const* q, C const* a, C const* b)                                                   \
    {                                                                                                                  \
        auto const d = atan2_num(load3(q), load3(a), load3(b));                                                        \
        return d > 0 ? 1 : (d < 0 ? -1 : 0);                                                                           \
    }                                                                                                                  \
    int orc_project_into_tri_##SFX(C const* q, C const* a, C const* b, C const* c, double* w3)                         \
    {                                                                                                                  \
        w3[0] = w3[1] = w3[2] = 0.0;                                                                                   \
        return project_into_tri(load3(q), load3(a), load3(b), load3(c), w3);                                           \
    }

ORC_PRED(int32_t, i32)
ORC_PRED(int64_t, i64)
ORC_PRED(double, f64)

// Batched predicate evaluation (n cases, packed q|a|b|c as 12 coordinates each) for large random sweeps.
void orc_pred_batch_f64(double const* cases, int64_t n, int32_t* edge_sign_out, int32_t* num_sign_out,
                        int32_t* inside_out, double* w_out)
{
    for (int64_t i = 0; i < n; ++i)
    {
        double const* c = cases + 12 * i;
        edge_sign_out[i] = edge_sign(load3(c), load3(c + 3), load3(c + 6));
        auto const nm = atan2_num(load3(c), load3(c + 3), load3(c + 6));
        num_sign_out[i] = nm > 0 ? 1 : (nm < 0 ? -1 : 0);
        inside_out[i] = project_into_tri(load3(c), load3(c + 3), load3(c + 6), load3(c + 9), w_out + 3 * i);
    }
}
void orc_pred_batch_i64(int64_t const* cases, int64_t n, int32_t* edge_sign_out, int32_t* num_sign_out,
                        int32_t* inside_out, double* w_out)
{
    for (int64_t i = 0; i < n; ++i)
    {
        int64_t const* c = cases + 12 * i;
        edge_sign_out[i] = edge_sign(load3(c), load3(c + 3), load3(c + 6));
        auto const nm = atan2_num(load3(c), load3(c + 3), load3(c + 6));
        num_sign_out[i] = nm > 0 ? 1 : (nm < 0 ? -1 : 0);
        inside_out[i] = project_into_tri(load3(c), load3(c + 3), load3(c + 6), load3(c + 9), w_out + 3 * i);
    }
}
void orc_pred_batch_i32(int32_t const* cases, int64_t n, int32_t* edge_sign_out, int32_t* num_sign_out,
                        int32_t* inside_out, double* w_out)
{
    for (int64_t i = 0; i < n; ++i)
    {
        int32_t const* c = cases + 12 * i;
        edge_sign_out[i] = edge_sign(load3(c), load3(c + 3), load3(c + 6));
        auto const nm = atan2_num(load3(c), load3(c + 3), load3(c + 6));
        num_sign_out[i] = nm > 0 ? 1 : (nm < 0 ? -1 : 0);
        inside_out[i] = project_into_tri(load3(c), load3(c + 3), load3(c + 6), load3(c + 9), w_out + 3 * i);
    }
}

// --- robust evaluator ------------------------------------------------------------------------------------------------
// scalar: 4 (float) or 8 (double); policy: 0 int64_grid, 1 int128_grid, 2 exact_float. Any output may be null.
int orc_robust_eval(int scalar, int policy, void const* verts, int64_t nv, int32_t const* idx, int64_t nidx,
                    void const* seg7, int64_t nseg, void const* pos, int64_t n, void* out_frac, int32_t* out_count,
                    void* out_eval, int nth)
{
    int const th = nt(nth);
    if (scalar == 4)
    {
        using T = float;
        auto const* v = static_cast<T const*>(verts);
        auto const* s = static_cast<T const*>(seg7);
        auto const* p = static_cast<T const*>(pos);
        if (policy == 0)
            robust_run<int32_t, T>(20, v, nv, idx, nidx, s, nseg, p, n, (T*)out_frac, out_count, (T*)out_eval, th);
        else if (policy == 1)
            robust_run<int64_t, T>(50, v, nv, idx, nidx, s, nseg, p, n, (T*)out_frac, out_count, (T*)out_eval, th);
        else if (policy == 2)
            robust_run<double, T>(0, v, nv, idx, nidx, s, nseg, p, n, (T*)out_frac, out_count, (T*)out_eval, th);
        else
            return -1;
        return 0;
    }
    if (scalar == 8)
    {
        using T = double;
        auto const* v = static_cast<T const*>(verts);
        auto const* s = static_cast<T const*>(seg7);
        auto const* p = static_cast<T const*>(pos);
        if (policy == 0)
            robust_run<int32_t, T>(20, v, nv, idx, nidx, s, nseg, p, n, (T*)out_frac, out_count, (T*)out_eval, th);
        else if (policy == 1)
            robust_run<int64_t, T>(50, v, nv, idx, nidx, s, nseg, p, n, (T*)out_frac, out_count, (T*)out_eval, th);
        else if (policy == 2)
            robust_run<double, T>(0, v, nv, idx, nidx, s, nseg, p, n, (T*)out_frac, out_count, (T*)out_eval, th);
        else
            return -1;
        return 0;
    }
    return -1;
}

// Quantization frame of the grid policies (for host-side checks of the device build): center xyz, scale.
void orc_robust_frame_f64(int grid_bits, double const* verts, int64_t nv, double* out4)
{
    auto const m = robust_build<int64_t, double>(grid_bits, verts, nv, nullptr, 0, nullptr, 0);
    out4[0] = m.cx;
    out4[1] = m.cy;
    out4[2] = m.cz;
    out4[3] = m.scale;
}

} // extern "C"
