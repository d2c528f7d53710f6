// Internal declarations shared by the HIP translation units of libagwn.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

struct agwn_mlp; // surrogate network handle (agwn_mlp.hip)

namespace agwn
{
// ---------------------------------------------------------------------------------------------------------------------
// error plumbing: every C entry point returns an int status; the message is kept per host thread.
// ---------------------------------------------------------------------------------------------------------------------
enum status : int
{
    ok = 0,
    err_invalid = 1,
    err_hip = 2,
    err_nodevice = 3,
    err_alloc = 4,
    err_unsupported = 5,
};

void set_error(char const* fmt, ...);
void clear_error();

struct failure
{
    int code;
};

// Throw-free inside kernels' launchers: these helpers record the message and return the status to the caller.
#define AGWN_HIP(call)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t const e__ = (call);                                                                                 \
        if (e__ != hipSuccess)                                                                                         \
        {                                                                                                              \
            ::agwn::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e__), __FILE__, __LINE__);              \
            return (e__ == hipErrorOutOfMemory || e__ == hipErrorMemoryAllocation) ? ::agwn::err_alloc                 \
                                                                                   : ::agwn::err_hip;                  \
        }                                                                                                              \
    } while (0)

#define AGWN_CHECK(cond, ...)                                                                                          \
    do                                                                                                                 \
    {                                                                                                                  \
        if (!(cond))                                                                                                   \
        {                                                                                                              \
            ::agwn::set_error(__VA_ARGS__);                                                                            \
            return ::agwn::err_invalid;                                                                                \
        }                                                                                                              \
    } while (0)

#define AGWN_TRY(expr)                                                                                                 \
    do                                                                                                                 \
    {                                                                                                                  \
        int const s__ = (expr);                                                                                        \
        if (s__ != ::agwn::ok)                                                                                         \
            return s__;                                                                                                \
    } while (0)

// Stream-ordered device buffer (hipMallocAsync / hipFreeAsync on the owning stream).
struct dbuf
{
    void* p = nullptr;
    hipStream_t s = nullptr;
    dbuf() = default;
    dbuf(dbuf const&) = delete;
    dbuf& operator=(dbuf const&) = delete;
    ~dbuf()
    {
        if (p)
            (void)hipFreeAsync(p, s);
    }
    // (re)allocates; a previous allocation is released first (its last use must be complete)
    int alloc(size_t bytes, hipStream_t stream)
    {
        if (p)
        {
            (void)hipFreeAsync(p, s);
            p = nullptr;
        }
        s = stream;
        if (bytes == 0)
            bytes = 256;
        AGWN_HIP(hipMallocAsync(&p, bytes, stream));
        return ok;
    }
    template <class U>
    U* as() const
    {
        return static_cast<U*>(p);
    }
};

// Device workspace that a handle keeps between calls: take() hands out 256-byte aligned pieces of a few large blocks,
// which are hipMalloc'd only when a call needs more than the blocks hold and freed with the arena; reset() makes all
// of it available again. A piece stays valid until the next reset(), so a call that resets must first have waited for
// the previous call's work on the pieces; calls using one arena are serialised by its owner. (Per-call stream-ordered
// allocation and release cost ~3 ms per labelling call of the 512^3 grid in the frees alone.)
struct arena
{
    struct block
    {
        char* p;
        size_t size, used;
    };
    std::vector<block> blocks;
    std::vector<int64_t> host; // host staging whose copy to the device must outlive the call (until the next reset)
    arena() = default;
    arena(arena const&) = delete;
    arena& operator=(arena const&) = delete;
    ~arena()
    {
        for (block& b : blocks)
            (void)hipFree(b.p);
    }
    void reset()
    {
        for (block& b : blocks)
            b.used = 0;
    }
    int take(size_t bytes, void** out)
    {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        for (block& b : blocks)
            if (b.size - b.used >= bytes)
            {
                *out = b.p + b.used;
                b.used += bytes;
                return ok;
            }
        size_t const size = std::max(bytes, blocks.empty() ? size_t(64) << 20 : 2 * blocks.back().size);
        void* p = nullptr;
        AGWN_HIP(hipMalloc(&p, size));
        blocks.push_back(block{static_cast<char*>(p), size, bytes});
        *out = p;
        return ok;
    }
    template <class U>
    int take_n(size_t count, U** out)
    {
        void* p = nullptr;
        AGWN_TRY(take(count * sizeof(U), &p));
        *out = static_cast<U*>(p);
        return ok;
    }
};

int ensure_device(); // selects nothing; checks a device is present (cached)

// ---------------------------------------------------------------------------------------------------------------------
// Rotated frame of the fractional term: row 0 is x1 = -x0 (the antipodal pole), rows 1, 2 complete a right-handed
// orthonormal frame. In that frame x1 . v = v.x and x1 . (v0 x v1) = v0.y v1.z - v0.z v1.y.
// ---------------------------------------------------------------------------------------------------------------------
struct frame
{
    double r[9];
};
frame make_pole_frame(double x0x, double x0y, double x0z);

// The FP32 kernels' coordinate scale: e such that the boundary's extent (the largest of its endpoints' per-axis
// max - min, in double) lies in [2^e, 2^(e+1)); 0 for a boundary without extent. Coordinates are multiplied by 2^-e
// (exact), so the FP32 pair's range limits are relative to the boundary's size and a power-of-two scaling of every
// input leaves each FP32 operation's rounding, and so every result, unchanged. far (may be null) receives the centre
// of the endpoints' box (caller frame) and 2^(26 + e): a query farther than that from the centre in any coordinate
// runs all its pairs through the FP64 path (the FP32 products of its pairs could overflow).
int boundary_scale_exp(float const* seg7_host, int64_t nseg, float4* far);
// R with every entry multiplied by 2^-e
frame scale_frame(frame const& R, int e);

// Packed, rotated boundary segment: endpoints a, b in the pole frame as float pairs (hi = the FP64 coordinate rounded
// to float, lo = the rounded remainder; hi + lo carries ~48 bits), weight w. 64 B.
struct alignas(64) seg_f32
{
    float ax, ay, az, bx, by, bz, w, pad0;
    float lax, lay, laz, lbx, lby, lbz, pad1, pad2;
};
// The same segment in double, for the FP64 path and the FP32 kernel's ill-conditioned-pair fallback. 64 B.
struct alignas(64) seg_f64
{
    double ax, ay, az, bx, by, bz, w, pad;
};

// ---------------------------------------------------------------------------------------------------------------------
// Query sources: a device array of positions, or a lattice generated in-kernel.
// ---------------------------------------------------------------------------------------------------------------------
template <class T>
struct src_array
{
    T const* pos;
};

// p(ix,iy,iz) = start + dx*ix + dy*iy + dz*iz in float, evaluated left to right per component (the caller-side
// positions of gwn_mesh_opengl.hh:511-519), flattened in ((iz*ny)+iy)*nx+ix order.
struct grid_desc
{
    float start[3], dx[3], dy[3], dz[3];
    int64_t nx, ny, nz;
    double inx, iny; // 1 / nx, 1 / ny (the lattice index is split with them; nx * ny * nz <= 2^52)
};

// q = a / d, r = a % d for 0 <= a < 2^52 and 0 < d < 2^31, with inv = 1 / d rounded to double: a * inv is within a
// relative 2^-52 of a / d, so its truncation is the quotient or one off it, and one correction step makes it exact.
// (The 64-bit integer division it replaces is a long emulated sequence on this GPU.)
__device__ __forceinline__ int64_t div_recip(int64_t a, int64_t d, double inv, int64_t& r)
{
    int64_t q = int64_t(double(a) * inv);
    r = a - q * d;
    if (r < 0)
    {
        --q;
        r += d;
    }
    else if (r >= d)
    {
        ++q;
        r -= d;
    }
    return q;
}
struct src_grid
{
    grid_desc g;
};

template <class T>
__device__ __forceinline__ void query_at(src_array<T> const& s, int64_t i, T& x, T& y, T& z)
{
    x = s.pos[3 * i];
    y = s.pos[3 * i + 1];
    z = s.pos[3 * i + 2];
}

// lattice cell (ix, iy, iz) of flat index i
__device__ __forceinline__ void grid_cell(grid_desc const& g, int64_t i, int64_t& ix, int64_t& iy, int64_t& iz)
{
    int64_t const t = div_recip(i, g.nx, g.inx, ix);
    iz = div_recip(t, g.ny, g.iny, iy);
}

__device__ __forceinline__ void query_at(src_grid const& s, int64_t i, float& x, float& y, float& z)
{
#pragma clang fp contract(off)
    int64_t ix, iy, iz;
    grid_cell(s.g, i, ix, iy, iz);
    float const fx = float(ix), fy = float(iy), fz = float(iz);
    x = ((s.g.start[0] + s.g.dx[0] * fx) + s.g.dy[0] * fy) + s.g.dz[0] * fz;
    y = ((s.g.start[1] + s.g.dx[1] * fx) + s.g.dy[1] * fy) + s.g.dz[1] * fz;
    z = ((s.g.start[2] + s.g.dx[2] * fx) + s.g.dy[2] * fy) + s.g.dz[2] * fz;
}

// Boundary prepared for the FP32 kernel's chain path (agwn_frac.hip): when every weight is a small integer, the
// segments are expanded to unit weight (endpoints swapped where the weight is negative) and ordered into chains, and
// stored as a stream of points (x, y, z, flag) in the caller's frame -- flag 1: a unit segment runs from the previous
// point to this one; flag 0: a chain starts here. Consecutive segments then share an endpoint. chain == false: the
// weights are general and the per-segment kernel runs instead.
//
// Every walk of a boundary that is a 1-cycle (build_boundary_segments' always is: the boundary of a triangle chain)
// ends where it starts, so the walks are closed loops (long ones split into closed arcs plus a skeleton loop whose
// chords cancel the arcs' closing chords). For each, in double and the caller's frame: a bounding ball
// (centre c = mean of its points, radius rho), the vector area a = sum over the fan from c of (1/2) u x w, the area of
// that fan (a spanning surface inside the ball) and its first moment M_ij = sum of (triangle centroid - c)_i (n dA)_j.
// They bound the loop's fractional contribution at a query outside the ball (frac_decide). loops_ok is false when
// some walk is open (then there is no decision stage).
struct loop_host
{
    double c[3];
    double rho;
    double a[3];
    double area;
    double m[9];
    double inertia; // second moment of area about c: the integral of |y - c|^2 over the fan(s), <= rho^2 area
    double s[3];    // first moment of area about c: the integral of y - c (for moving the second moment)
    int64_t start, count; // decision-stream entries [start, start + count)
};
struct frac_plan
{
    dbuf buf;           // npts float4 entries (device)
    int64_t npts = 0;   // stream entries
    int64_t nunit = 0;  // unit segments
    int sexp = 0;       // boundary_scale_exp of the boundary: the FP32 kernels' frame is scaled by 2^-sexp
    float4 far{0.0f, 0.0f, 0.0f, HUGE_VALF}; // its far-query test (centre, threshold)
    bool chain = false;
    std::vector<int64_t> walk_starts; // stream entries with flag 0 after an even number of segments (host copy;
                                      // the cuts of small launches)
    std::vector<loop_host> loops; // closed loops (long walks split into arcs + a skeleton), in the decision stream
    bool loops_ok = false;
    dbuf dec;                     // decision stream: the loops' points, each loop contiguous (flag 0 at its start)
    std::vector<float4> dec_host; // its host copy (the loop tree's node balls are fitted to the points)
    int64_t dec_npts = 0;
    // the decision's loop tree and, per pole, its device tables: built by the plan's first decision for that pole
    // (under dec_mu) and kept with the plan, read-only afterwards (agwn_frac.hip)
    mutable std::mutex dec_mu;
    mutable std::shared_ptr<void> dec_state;
    float4 const* pts() const { return buf.as<float4 const>(); }
};
// Builds the plan from a host copy of the boundary (stream-ordered upload on `stream`; returns when it is done).
int frac_plan_build(float const* seg7_host, int64_t nseg, frac_plan& plan, hipStream_t stream);

// A boundary's plan kept for later calls: reused when the device copy of the boundary it was built from equals the
// call's boundary word for word (compared on the device), rebuilt otherwise. Calls are serialised by mu; a plan handed
// out stays alive with its holders when a later call replaces it.
struct plan_cache
{
    std::mutex mu;
    std::shared_ptr<frac_plan> plan;
    float* seg = nullptr; // device copy of the boundary the plan was built from
    int64_t nseg = -1;
    int dev = -1;
    plan_cache() = default;
    plan_cache(plan_cache const&) = delete;
    plan_cache& operator=(plan_cache const&) = delete;
    ~plan_cache()
    {
        if (seg)
            (void)hipFree(seg);
    }
};
int cached_plan(plan_cache& c, float const* seg7, int64_t nseg, hipStream_t s, std::shared_ptr<frac_plan const>& out);

// Inside/outside decision without the full boundary sum, for m query points (device, FP32 xyz) with their integer
// terms `counts` (needs plan.chain && plan.loops_ok): per query, loops whose bound is small are replaced by a
// second-order far-field expansion with a rigorous remainder bound, the others are summed exactly (FP32 chain
// arithmetic, FP64 for ill-conditioned pairs). Where count + estimate is farther from 0.5 than the bound plus a margin
// that covers the FP32 evaluation's own error, val[i] = that estimate (on the same side of 0.5 as the full GWN);
// otherwise i is appended to `undecided` (device, m entries) and val[i] is left unwritten. nund: how many (host).
// beta: loops whose remainder bound (in units of the fractional term) exceeds it are summed exactly -- smaller is
// tighter and slower; labelling runs a first pass at kDecideBeta and a second at kDecideBetaTight over what it leaves.
constexpr float kDecideBeta = 0.002f;
constexpr float kDecideBetaTight = 0.0002f;
int frac_decide(float const* pos, int64_t m, int32_t const* counts, frac_plan const& plan, double const x0[3],
                float beta, float* val, int64_t* undecided, int64_t& nund, hipStream_t s, arena* ws = nullptr);

// Fractional launchers (agwn_frac.hip). seg7/out/counts are device pointers. counts may be null (fractional
// only); otherwise out[i] = fractional + counts[i] (the full GWN of gwn_mesh.hh:213-218).
// Instantiated for src_array<float>, src_grid (FP32) and src_array<double> (FP64).
template <class Src>
int frac_f32_src(Src const& src, int64_t n, float const* seg7, int64_t nseg, double const x0[3],
                 int32_t const* counts, float* out, hipStream_t stream);
// chain: a boundary of small integer weights is walked as a double chain stream (each stream point evaluated once,
// paired arctangents: the full-GWN path); otherwise, and for general weights, segment by segment
template <class Src>
int frac_f64_src(Src const& src, int64_t n, double const* seg7, int64_t nseg, double const x0[3],
                 int32_t const* counts, double* out, hipStream_t stream, bool chain = false);
// The FP64 launcher with the caller's workspace: ws holds max(nseg, 1) seg_f64 (device), the boundary rotated into
// the pole frame of x0. Launches two kernels on `stream`; allocates nothing. Instantiated for src_array<double>.
template <class Src>
int frac_f64_ws(Src const& src, int64_t n, double const* seg7, int64_t nseg, double const x0[3],
                int32_t const* counts, double* out, seg_f64* ws, hipStream_t stream);

// Surrogate routing (agwn_mlp.hip), instantiated for src_array<float> and src_grid:
//   mlp_route      network labels of n points -> labels; the band / near-surface / audit points are appended to list
//                  (2*i + audit) and counted in counters[0] (band), [1] (near surface or outside the box), [2] (audit),
//                  [3] (list length); counters (5 of them) start zeroed
//   gather_routed  pts[e] = the query of list entry e (the coordinates the exact kernels would compute)
//   route_apply    labels of list entries from their exact values w; audit disagreements counted in counters[4]
// lat: for a lattice source, device room for mlp_lattice_table_floats(g) floats (null, or a lattice whose layer 0 does
// not split by axis: the per-point features); ignored for arrays
template <class Src>
int mlp_route(agwn_mlp const* mlp, Src const& src, int64_t n, float delta, uint32_t audit_mod, uint32_t audit_seed,
              int8_t* labels, int64_t* list, unsigned long long* counters, hipStream_t s, float* lat = nullptr);
// floats of the layer-0 table of an axis-separable lattice (0: the lattice's layer 0 does not split by axis, or its
// table would exceed 2^16 rows)
int64_t mlp_lattice_table_floats(grid_desc const& g);
// network outputs of the n = nx ny nz lattice points, in the lattice's flat order
int mlp_predict_grid_dev(agwn_mlp const* mlp, grid_desc const& g, float* out, hipStream_t s);
template <class Src>
int gather_routed(Src const& src, int64_t const* list, int64_t m, float* pts, hipStream_t s);
//   route_order    orders the m list entries of n queries: *by_index = the entries in index order (list or tmp);
//                  with perm, *perm[e] = the position in *by_index of the e-th entry in the Morton order of the
//                  queries in the network's box (null when position and code would not fit 64 bits: index order);
//                  tmp, keys, keys2: m uint64 of scratch each; sort_ws: radix_sort_words(m) uint32 of scratch
template <class Src>
int route_order(agwn_mlp const* mlp, Src const& src, int64_t n, int64_t* list, int64_t m, uint64_t* tmp,
                uint64_t* keys, uint64_t* keys2, uint32_t* sort_ws, hipStream_t s, int64_t const** by_index,
                uint64_t const** perm);
// LSD radix sort of n uint64 keys over their low `bits` bits (agwn_boundary.hip); *sorted = a or b. The _ws form takes
// its scratch (radix_sort_words(n) uint32) from the caller and only launches kernels (graph capture).
int radix_sort_u64(uint64_t* a, uint64_t* b, int64_t n, int bits, hipStream_t s, uint64_t** sorted);
int64_t radix_sort_words(int64_t n);
int radix_sort_u64_ws(uint64_t* a, uint64_t* b, int64_t n, int bits, uint32_t* ws, hipStream_t s, uint64_t** sorted);
int route_apply(int64_t const* list, int64_t m, float const* w, int8_t* labels, unsigned long long* counters,
                hipStream_t s);

// Batched evaluation (agwn_frac.hip): the persistent chain kernel over many meshes' point streams.
struct batch_frac_desc
{
    int64_t pt0, npts; // the mesh's point stream within the concatenated stream
    int64_t q_end;     // one past the mesh's last query in the concatenated query array
    double qscale;     // 2^-e, e = boundary_scale_exp of the mesh's boundary (its queries are scaled by it)
    float4 far;        // the boundary's far-query test (boundary_scale_exp)
};
bool chain_stream(float const* seg7_host, int64_t nseg, std::vector<float4>& stream_pts, int64_t& nunit);
int frac_batch_variants(); // number of (queries per lane, block, LDS stage) shapes
// shape `variant`: q queries per lane, `block` lanes, stage x block stream points per LDS stage
void frac_batch_shape(int variant, int& q, int& block, int& stage);
int frac_batch_blocks_per_cu(int variant); // resident workgroups per CU (wave and LDS limits)
// rotates the concatenated stream into the pole frame of x0, each point scaled by 2^-pexp[i] (its mesh's exponent;
// device int16 per point): two float4 per point, (hi, flag) and (lo, 0)
int frac_stream_rotate(float4 const* in, int64_t n, double const x0[3], int16_t const* pexp, float4* out,
                       hipStream_t s);
int frac_batch_launch(int variant, int blocks, float const* pos, batch_frac_desc const* meshes,
                      int32_t const* item_mesh, int64_t const* item_q0, int64_t nitems, unsigned long long* counter,
                      float4 const* pts, double const x0[3], int32_t const* counts, float* out, hipStream_t s);

// The FP32 launcher with a prepared plan (seg7 is still needed when the plan is not a chain plan).
template <class Src>
// split_small: a launch with few query tiles may split the stream into chunks summed separately (fills the GPU; the
// FP64 sums then add in a different order, so values may differ from the unsplit kernel's in the last float bit --
// only for callers that need the side of 0.5, not bit-stable values)
int frac_f32_planned(Src const& src, int64_t n, frac_plan const& plan, float const* seg7, int64_t nseg,
                     double const x0[3], int32_t const* counts, float* out, hipStream_t stream,
                     bool split_small = false, arena* ws = nullptr);

inline int frac_f32(float const* pos, int64_t n, float const* seg7, int64_t nseg, double const x0[3],
                    int32_t const* counts, float* out, hipStream_t stream)
{
    return frac_f32_src(src_array<float>{pos}, n, seg7, nseg, x0, counts, out, stream);
}
inline int frac_f64(double const* pos, int64_t n, double const* seg7, int64_t nseg, double const x0[3],
                    int32_t const* counts, double* out, hipStream_t stream, bool chain = false)
{
    return frac_f64_src(src_array<double>{pos}, n, seg7, nseg, x0, counts, out, stream, chain);
}
} // namespace agwn
