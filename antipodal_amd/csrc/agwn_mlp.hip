// Learned inside/outside surrogate: a per-mesh Fourier-feature MLP fitted to exact GWN values, with hand-written BF16
// MFMA kernels (FP32 accumulation) for the forward pass, the backward pass and Adam, and a fused inference kernel that
// labels points and routes the uncertain ones to the exact path.
//
// Network (fixed shape):
//   u_a = (p_a - c_a) / h_a                     c, h: centre and half-extent of the training box, per axis
//   features f[2j], f[2j+1] = sin(pi w_k u_a), cos(pi w_k u_a),  j = 32 a + k,  w_k = F^(k/31)   (192 features;
//                                                                               F = max_freq, 128 by default)
//   z_l = H_{l-1} W_l + b_l,  H_l = hsilu(z_l)  l = 0..3, 256 units, H_{-1} = f,  hsilu(z) = z clamp(z/4 + 1/2, 0, 1)
//   pred = H_3 . wo + bo
// Rounding points (what the kernels store, and what the numerics tests emulate): features, weights W_l, activations
// H_l and backward activations dz_l are BF16; products accumulate in FP32; z_l, biases, wo, bo, the loss and every
// gradient are FP32.
//
// MFMA: v_mfma_f32_32x32x16_bf16. Lane l holds A[row = l%32][k = 8(l/32) .. +8] and B[k = 8(l/32) .. +8][col = l%32];
// accumulator e of lane l is C[8(e/4) + 4(l/32) + e%4][l%32]. Every GEMM here is written as C[r][c] = sum_k A[r][k]
// B[c][k] with both operands k-contiguous ("row" arrays), so each lane loads 16 contiguous bytes per operand.
//
// Parameter vector (FP32 master copy, in this order): W0 [192 x 256], b0 [256], W1..W3 [256 x 256] with their biases,
// wo [256], bo. W_l[k][n] maps input k to unit n. BF16 copies Wr_l (= W_l, row-major) and Wt_l ([n][k]) feed MFMA.

#include "../../include/agwn.h"
#include "agwn_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

namespace agwn
{
namespace
{
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kFreq = 32;
constexpr int kIn = 3 * kFreq * 2; // 192
constexpr int kH = 256;
constexpr int kL = 4; // hidden layers
constexpr int kB = 256;

constexpr int64_t kInDim[kL] = {kIn, kH, kH, kH};

struct layout
{
    int64_t w[kL], b[kL], wo, bo, total;
    constexpr layout() : w{}, b{}, wo(0), bo(0), total(0)
    {
        int64_t o = 0;
        for (int l = 0; l < kL; ++l)
        {
            w[l] = o;
            o += kInDim[l] * kH;
            b[l] = o;
            o += kH;
        }
        wo = o;
        o += kH;
        bo = o;
        o += 1;
        total = o;
    }
};
constexpr layout kLay{};

// feature map of one MLP (the training box and the frequencies)
struct feat_params
{
    float c[3], inv_h[3];
    float w[kFreq];
};

// The activation: hard SiLU, z clamp(z/4 + 1/2, 0, 1) -- SiLU with its sigmoid replaced by the sigmoid's tangent at 0,
// clamped -- and its derivative, z/2 + 1/2 for |z| < 2, 1 above, 0 below (the one-sided values at z = +-2 are 1 and
// 0). Three VALU instructions per unit where SiLU's exponential and reciprocal cost a fifth of the fused inference
// kernel (1024 activations per point); networks trained with it route fewer points than SiLU's on the benchmark mesh.
// Training and inference share these definitions.
__device__ __forceinline__ float hsilu(float z) { return z * __builtin_amdgcn_fmed3f(fmaf(0.25f, z, 0.5f), 0.0f, 1.0f); }
__device__ __forceinline__ float hsilu_grad(float z)
{
    return fabsf(z) < 2.0f ? fmaf(0.5f, z, 0.5f) : (z >= 2.0f ? 1.0f : 0.0f);
}

// feature pair j (0..95) of a point: axis a = j / 32, frequency k = j % 32: (sin, cos)(pi w_k u), u the box-normalised
// coordinate. The hardware sine and cosine take revolutions (sin(2 pi t)), so t = w_k u / 2; their error (~1e-6 over
// |t| <= 256, here |t| <= F / 2 for points in the box) is far below the BF16 rounding the features go through, and a
// full-precision sincospi cost as much as a quarter of the fused inference kernel. Training and inference share it.
__device__ __forceinline__ void feature_pair(feat_params const& F, float x, float y, float z, int j, float& s, float& c)
{
    int const a = j / kFreq, k = j % kFreq;
    float const p = a == 0 ? x : (a == 1 ? y : z);
    float const t = 0.5f * (F.w[k] * ((p - F.c[a]) * F.inv_h[a]));
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
}

// Training features: row m of X [rows x 192] and column m of XT [192 x rows] for sample perm[m] (m < nvalid; rows past
// nvalid are zero).
__global__ void __launch_bounds__(kB) k_features(float const* __restrict__ pts, int64_t const* __restrict__ perm,
                                                 int64_t nvalid, int64_t rows, feat_params F, __bf16* __restrict__ X,
                                                 __bf16* __restrict__ XT)
{
    int64_t const t = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (t >= rows * (kIn / 2))
        return;
    // consecutive threads take consecutive samples of one feature pair: the XT stores coalesce
    int const j = int(t / rows);
    int64_t const m = t % rows;
    float s = 0.0f, c = 0.0f;
    if (m < nvalid)
    {
        int64_t const src = perm ? perm[m] : m;
        feature_pair(F, pts[3 * src], pts[3 * src + 1], pts[3 * src + 2], j, s, c);
    }
    __bf16 const bs = __bf16(s), bc = __bf16(c);
    X[m * kIn + 2 * j] = bs;
    X[m * kIn + 2 * j + 1] = bc;
    XT[int64_t(2 * j) * rows + m] = bs;
    XT[int64_t(2 * j + 1) * rows + m] = bc;
}

// ---------------------------------------------------------------------------------------------------------------------
// Training GEMM: C[r][c] = sum_{k in chunk} A[r*lda + k] * B[c*ldb + k], r < R, c < Cn. Block tile 128 x 128 (4 waves,
// 2 x 2, each 64 x 64 = 2 x 2 MFMA blocks; operands staged through LDS 256 k at a time); R, Cn multiples of 64, K
// chunk a multiple of 16. blockIdx.x tiles c,
// blockIdx.y tiles r, blockIdx.z splits K (partial sums, EPI_PART). The epilogue gets each lane's runs of 4
// consecutive rows r..r+3 at one column c.
// ---------------------------------------------------------------------------------------------------------------------
enum epi_kind
{
    EPI_FWD,  // r = unit n, c = sample m: z = acc + bias[n] -> Z[m][n] (FP32), H[m][n], HT[n][m] = bf16(hsilu(z))
    EPI_DX,   // r = input k, c = sample m: dz = acc * hsilu'(Z[m][k]) -> D[m][k], DT[k][m] (BF16)
    EPI_PART, // r = input k, c = unit n: P[blockIdx.z][k][n] = acc
};

struct epi_args
{
    float const* bias;
    float const* zin; // EPI_DX: pre-activations of the layer below [rows x R]
    float* zout;      // EPI_FWD
    __bf16* h;        // EPI_FWD: H / EPI_DX: D, [rows x R]
    __bf16* ht;       // transposed copy [R x rows]
    float* part;      // EPI_PART
    int64_t rows;     // samples in the batch (EPI_FWD / EPI_DX)
};

template <int EPI>
__global__ void __launch_bounds__(256)
    k_gemm(__bf16 const* __restrict__ A, int64_t lda, __bf16 const* __restrict__ B, int64_t ldb, int64_t R, int64_t Cn,
           int64_t K, int64_t kchunk, epi_args E)
{
    // The block's 128 rows of A and 128 rows of B are staged in LDS 256 k at a time (each thread keeps sixteen 16-byte
    // loads in flight: two round trips per stage), then the waves' MFMAs read their operands from LDS. 528-byte LDS
    // rows keep the 16-lane groups of ds_read_b128 conflict-free.
    constexpr int kS = 256 + 8;
    __shared__ __attribute__((aligned(16))) __bf16 sa[128 * kS];
    __shared__ __attribute__((aligned(16))) __bf16 sb[128 * kS];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int64_t const rb = int64_t(blockIdx.y) * 128, cb = int64_t(blockIdx.x) * 128;
    int64_t const r0 = rb + (wave >> 1) * 64;
    int64_t const c0 = cb + (wave & 1) * 64;
    bool const active = r0 < R && c0 < Cn;
    int64_t const kb = int64_t(blockIdx.z) * kchunk;
    int64_t const ke = kb + kchunk < K ? kb + kchunk : K;
    int const li = lane & 31;
    int const kh = (lane >> 5) * 8;
    f32x16 acc[2][2] = {};
    for (int64_t k1 = kb; k1 < ke; k1 += 256)
    {
        int const kn = int(ke - k1 < 256 ? ke - k1 : 256); // a multiple of 16
        int const cpr = kn / 8;                           // 16-byte chunks per row
        int const nch = 128 * cpr;
        for (int q0 = threadIdx.x; q0 < nch; q0 += 8 * 256)
        {
            bf16x8 va[8], vb[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                int const q = q0 + 256 * u;
                int const row = q / cpr, ch = q - row * cpr;
                va[u] = bf16x8{};
                vb[u] = bf16x8{};
                if (q < nch && rb + row < R)
                    va[u] = *reinterpret_cast<bf16x8 const*>(A + (rb + row) * lda + k1 + 8 * ch);
                if (q < nch && cb + row < Cn)
                    vb[u] = *reinterpret_cast<bf16x8 const*>(B + (cb + row) * ldb + k1 + 8 * ch);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
            {
                int const q = q0 + 256 * u;
                int const row = q / cpr, ch = q - row * cpr;
                if (q < nch)
                {
                    *reinterpret_cast<bf16x8*>(sa + row * kS + 8 * ch) = va[u];
                    *reinterpret_cast<bf16x8*>(sb + row * kS + 8 * ch) = vb[u];
                }
            }
        }
        __syncthreads();
        if (active)
        {
            __bf16 const* la[2] = {sa + (r0 - rb + li) * kS + kh, sa + (r0 - rb + 32 + li) * kS + kh};
            __bf16 const* lb[2] = {sb + (c0 - cb + li) * kS + kh, sb + (c0 - cb + 32 + li) * kS + kh};
            for (int k0 = 0; k0 < kn; k0 += 16)
            {
                bf16x8 a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                {
                    a[i] = *reinterpret_cast<bf16x8 const*>(la[i] + k0);
                    b[i] = *reinterpret_cast<bf16x8 const*>(lb[i] + k0);
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads(); // the stage has been read before the next one overwrites it
    }
    if (!active)
        return;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                int64_t const r = r0 + i * 32 + 8 * q + 4 * (lane >> 5); // rows r .. r+3
                int64_t const c = c0 + j * 32 + li;
                float v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    v[e] = acc[i][j][4 * q + e];
                if constexpr (EPI == EPI_FWD)
                {
                    float z[4];
                    bf16x4 hv;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                    {
                        z[e] = v[e] + E.bias[r + e];
                        hv[e] = __bf16(hsilu(z[e]));
                    }
                    *reinterpret_cast<float4*>(E.zout + c * R + r) = make_float4(z[0], z[1], z[2], z[3]);
                    *reinterpret_cast<bf16x4*>(E.h + c * R + r) = hv;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        E.ht[(r + e) * E.rows + c] = hv[e];
                }
                else if constexpr (EPI == EPI_DX)
                {
                    float4 const zz = *reinterpret_cast<float4 const*>(E.zin + c * R + r);
                    float const zv[4] = {zz.x, zz.y, zz.z, zz.w};
                    bf16x4 dv;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        dv[e] = __bf16(v[e] * hsilu_grad(zv[e]));
                    *reinterpret_cast<bf16x4*>(E.h + c * R + r) = dv;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        E.ht[(r + e) * E.rows + c] = dv[e];
                }
                else
                {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        E.part[(int64_t(blockIdx.z) * R + r + e) * Cn + c] = v[e];
                }
            }
}

// out[i] = sum over p of part[p * count + i], in a fixed order (deterministic); the loads of eight partials are issued
// before their sums
__global__ void __launch_bounds__(kB) k_sum_parts(float const* __restrict__ part, int64_t parts, int64_t count,
                                                  float* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= count)
        return;
    float s = 0.0f;
    int64_t p = 0;
    for (; p + 8 <= parts; p += 8)
    {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            v[k] = part[(p + k) * count + i];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            s += v[k];
    }
    for (; p < parts; ++p)
        s += part[p * count + i];
    out[i] = s;
}

// Bias gradient: out[n] = sum_m DT[n][m] (one block per unit n; rows is a multiple of 128: each thread sums 8
// consecutive values per 16-byte load, then the block reduces)
__global__ void __launch_bounds__(kB) k_rowsum(__bf16 const* __restrict__ DT, int64_t rows, float* __restrict__ out)
{
    __shared__ float red[kB / 64];
    __bf16 const* row = DT + int64_t(blockIdx.x) * rows;
    float s = 0.0f;
    for (int64_t m = 8 * int64_t(threadIdx.x); m < rows; m += 8 * kB)
    {
        bf16x8 const v = *reinterpret_cast<bf16x8 const*>(row + m);
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            t += float(v[k]);
        s += t;
    }
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

constexpr int64_t kOblRows = 128; // rows per block of the output-layer kernel (rows are a multiple of 128)

// Output layer and loss over kOblRows rows per block. Wave-per-sample; lane l covers units 4l .. 4l+3.
//   pred = H3 . wo + bo;  g = 2 (pred - y) / nvalid;  dz3 = bf16(g wo hsilu'(Z3));  per-block partials of g H3 (dwo),
//   g (dbo) and (pred - y)^2.
// D is written row by row; its transpose DT goes through an LDS tile of the block's rows and is written unit by unit,
// 16 lanes per 256-byte unit row (whole cache lines). y is read through perm (sample order of the batch). Without D
// only pred is written.
__global__ void __launch_bounds__(kB)
    k_out_loss(__bf16 const* __restrict__ H, float const* __restrict__ Z, float const* __restrict__ wo,
               float const* __restrict__ bo_p, float const* __restrict__ y, int64_t const* __restrict__ perm,
               int64_t nvalid, int64_t rows, float* __restrict__ pred, __bf16* __restrict__ D, __bf16* __restrict__ DT,
               float* __restrict__ part_wo, float* __restrict__ part_sc)
{
    constexpr int kTs = kOblRows + 8; // LDS row stride of the transposed tile
    __shared__ __attribute__((aligned(16))) __bf16 dt[kH * kTs];
    __shared__ float sw[4][kH];
    __shared__ float ss[4][2];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int64_t const r0 = int64_t(blockIdx.x) * kOblRows;
    float w[4];
    for (int j = 0; j < 4; ++j)
        w[j] = wo[4 * lane + j];
    float const bo = *bo_p;
    float const inv = 1.0f / float(nvalid);
    float gw[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float gb = 0.0f, loss = 0.0f;
    // four rows per iteration (rows mr, mr + 4, mr + 8, mr + 12 of the wave's stride-4 share): their loads and
    // reductions are independent, the accumulations keep the row order
    constexpr int kU = 4;
    for (int mr0 = wave; mr0 < kOblRows; mr0 += 4 * kU)
    {
        float h[kU][4], s[kU], zv[kU][4], target[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u)
        {
            int64_t const m = r0 + mr0 + 4 * u;
            bf16x4 const hv = *reinterpret_cast<bf16x4 const*>(H + m * kH + 4 * lane);
            s[u] = 0.0f;
            for (int j = 0; j < 4; ++j)
            {
                h[u][j] = float(hv[j]);
                s[u] += h[u][j] * w[j];
            }
            if (D)
            {
                float4 const zz = *reinterpret_cast<float4 const*>(Z + m * kH + 4 * lane);
                zv[u][0] = zz.x;
                zv[u][1] = zz.y;
                zv[u][2] = zz.z;
                zv[u][3] = zz.w;
                target[u] = m < nvalid ? y[perm ? perm[m] : m] : 0.0f;
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int u = 0; u < kU; ++u)
                s[u] += __shfl_xor(s[u], off);
#pragma unroll
        for (int u = 0; u < kU; ++u)
        {
            int const mr = mr0 + 4 * u;
            int64_t const m = r0 + mr;
            float const p = s[u] + bo;
            bool const valid = m < nvalid;
            if (pred && lane == 0 && valid)
                pred[m] = p;
            if (!D)
                continue;
            float const tg = valid ? target[u] : p;
            float const g = valid ? 2.0f * (p - tg) * inv : 0.0f;
            bf16x4 dv;
            for (int j = 0; j < 4; ++j)
            {
                dv[j] = __bf16(g * w[j] * hsilu_grad(zv[u][j]));
                gw[j] += g * h[u][j];
                dt[(4 * lane + j) * kTs + mr] = dv[j];
            }
            *reinterpret_cast<bf16x4*>(D + m * kH + 4 * lane) = dv;
            gb += g;
            loss += valid ? (p - tg) * (p - tg) : 0.0f;
        }
    }
    if (!D)
        return;
    for (int j = 0; j < 4; ++j)
        sw[wave][4 * lane + j] = gw[j];
    if (lane == 0)
    {
        ss[wave][0] = gb;
        ss[wave][1] = loss;
    }
    __syncthreads();
    // DT[n][r0 .. r0 + 128): wave w writes units 64 w .. 64 w + 63, four units per store (16 lanes x 16 bytes each)
#pragma unroll 4
    for (int it = 0; it < 16; ++it)
    {
        int const n = 64 * wave + 4 * it + (lane >> 4);
        int const c = (lane & 15) * 8;
        *reinterpret_cast<bf16x8*>(DT + int64_t(n) * rows + r0 + c) = *reinterpret_cast<bf16x8 const*>(dt + n * kTs + c);
    }
    int const t = threadIdx.x;
    part_wo[int64_t(blockIdx.x) * kH + t] = (sw[0][t] + sw[1][t]) + (sw[2][t] + sw[3][t]);
    if (t < 2)
        part_sc[int64_t(blockIdx.x) * 2 + t] = (ss[0][t] + ss[1][t]) + (ss[2][t] + ss[3][t]);
}

// Reduces the output-layer partials of the kernel's nblk blocks (in block order): g_wo, g_bo, and adds the batch's
// squared-error sum to loss_acc (double).
__global__ void __launch_bounds__(kB) k_out_final(float const* __restrict__ part_wo, float const* __restrict__ part_sc,
                                                  int64_t nblk, float* __restrict__ g_wo, float* __restrict__ g_bo,
                                                  double* __restrict__ loss_acc)
{
    int const t = threadIdx.x;
    float s = 0.0f;
    int64_t p = 0;
    for (; p + 8 <= nblk; p += 8)
    {
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            v[k] = part_wo[(p + k) * kH + t];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            s += v[k];
    }
    for (; p < nblk; ++p)
        s += part_wo[p * kH + t];
    g_wo[t] = s;
    if (t < 64)
    {
        // lane t sums blocks t, t + 64, ... in order; the lanes' sums are then combined in a fixed tree
        float gb = 0.0f, l = 0.0f;
        for (int64_t q = t; q < nblk; q += 64)
        {
            gb += part_sc[2 * q];
            l += part_sc[2 * q + 1];
        }
        for (int off = 32; off > 0; off >>= 1)
        {
            gb += __shfl_xor(gb, off);
            l += __shfl_xor(l, off);
        }
        if (t == 0)
        {
            *g_bo = gb;
            if (loss_acc)
                *loss_acc += double(l);
        }
    }
}

struct wcopies
{
    __bf16* wr[kL];
    __bf16* wt[kL];
    __bf16* wf[kL];
};

// Inference copy of a hidden layer's weights in MFMA fragment order: the A operand of unit block ub (32 units) and k
// step st (16 inputs) is 64 lanes x 8 BF16 contiguous in lane order (lane L: unit 32 ub + L % 32, inputs
// 16 st + 8 (L / 32) + [0, 8)), so each wave's weight load is one coalesced 1 KiB read. (From Wt rows, a lane's 16 B
// came from a different 512 B row each: 32 cache lines per load, a quarter of each used -- the inference kernel's
// bound, ahead of the MFMAs.)
__host__ __device__ __forceinline__ int64_t frag_index(int64_t k, int64_t n, int64_t K)
{
    int64_t const lane = (n % 32) + 32 * ((k % 16) / 8);
    return (((n / 32) * (K / 16) + k / 16) * 64 + lane) * 8 + k % 8;
}

__device__ __forceinline__ void store_bf16_copies(wcopies const& C, int64_t i, float v)
{
#pragma unroll
    for (int l = 0; l < kL; ++l)
    {
        int64_t const o = kLay.w[l];
        int64_t const K = kInDim[l];
        if (i >= o && i < o + K * kH)
        {
            int64_t const k = (i - o) / kH, n = (i - o) % kH;
            __bf16 const b = __bf16(v);
            C.wr[l][k * kH + n] = b;
            C.wt[l][n * K + k] = b;
            C.wf[l][frag_index(k, n, K)] = b;
        }
    }
}

// Adam on the FP32 master copy (bias-corrected: c1 = 1/(1-b1^t), c2 = 1/(1-b2^t)), then the BF16 operand copies.
__global__ void __launch_bounds__(kB) k_adam(float* __restrict__ w, float* __restrict__ m1, float* __restrict__ m2,
                                             float const* __restrict__ g, int64_t n, float lr, float b1, float b2,
                                             float eps, float c1, float c2, wcopies C)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n)
        return;
    float const gi = g[i];
    float const a = b1 * m1[i] + (1.0f - b1) * gi;
    float const v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    float const wn = w[i] - lr * (a * c1) / (sqrtf(v * c2) + eps);
    w[i] = wn;
    store_bf16_copies(C, i, wn);
}

__global__ void __launch_bounds__(kB) k_cast(float const* __restrict__ w, int64_t n, wcopies C)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i < n)
        store_bf16_copies(C, i, w[i]);
}

// sample order of a batch: row j -> (offset + (b0 + j) * stride) mod n with stride coprime to n (a permutation per
// epoch)
__global__ void __launch_bounds__(kB) k_perm(int64_t* __restrict__ perm, int64_t rows, int64_t b0, int64_t offset,
                                             int64_t stride, int64_t n)
{
    int64_t const j = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (j >= rows)
        return;
    unsigned __int128 const t = (unsigned __int128)(uint64_t(b0 + j)) * uint64_t(stride) + uint64_t(offset);
    perm[j] = int64_t(uint64_t(t % uint64_t(n)));
}

// ---------------------------------------------------------------------------------------------------------------------
// Fused inference: a workgroup takes kTileP = 128 points through the features, the 4 hidden layers and the output
// layer with the activations in one LDS buffer (128 x 264 BF16, 66 KiB: two workgroups per CU; with weight fragments
// fetched three k steps ahead, 128-point tiles measured 4 % faster than 96-point tiles at three per CU -- each weight
// fragment feeds four MFMAs instead of three). Each layer computes
// out^T[n][m] = sum_k Wt[n][k] act[m][k]: wave w owns units [64w, 64w + 64) (A operand = Wt rows in fragment order,
// from L2) for all the tile's points (B operand = LDS rows); the accumulators start at the biases, so a lane's 4
// consecutive accumulators are 4 consecutive units of one point, activated and written back to LDS as one 8-byte
// write after a barrier. The output layer (a 256-long dot per point) is fused into the last layer's epilogue.
// Optionally labels the points (pred > 0.5) and appends to a routing list (entry = 2 * index + (audit ? 1 : 0)) the
// points in the near-surface mask or outside the box, those with |pred - 0.5| <= delta, and a hashed audit sample of
// the rest.
// ---------------------------------------------------------------------------------------------------------------------
constexpr int kLdsStride = kH + 8; // BF16 per LDS row: 528 B rows keep the 16-lane groups of ds_read_b128 conflict-free
constexpr int kTileP = 128;         // points per workgroup (a multiple of 32)
constexpr int kTileJ = kTileP / 32; // 32-point MFMA column blocks per tile
constexpr int kInferBlocks = (160 * 1024) / (kTileP * kLdsStride * 2 + 16 * kTileP + 4 * (kL + 1) * kH); // workgroups per CU that the LDS holds
static_assert(kTileP % 32 == 0 && kInferBlocks >= 1, "tile shape");

struct infer_params
{
    __bf16 const* wf[kL]; // fragment order (frag_index)
    float const* b[kL];
    float const* wo;
    float const* bo;
};

// Near-surface voxel mask over the training box: bit ((iz*ny)+iy)*nx+ix is set when some triangle comes within eps
// of that voxel (conservatively). Points in set voxels, and points outside the box, are not the network's to label.
struct surface_mask
{
    uint32_t const* bits; // null: no mask
    float lo[3], inv[3];  // voxel coordinate = (p - lo) * inv
    int32_t n[3];
};

struct route_params
{
    float delta;
    uint32_t audit_mod; // 0: no audit
    uint32_t audit_seed;
    int8_t* labels;
    int64_t* list;
    unsigned long long* counters; // [0] band, [1] near surface / outside the box, [2] audit, [3] list length,
                                  // [4] audit disagreements
    surface_mask mask;
};

// the near-surface test in two halves: the fetch of the mask word that holds the point's voxel (a point outside the box,
// or NaN, gets the word 1 and bit 0: near) and the bit test, so that a caller can issue the load early and test late
__device__ __forceinline__ void near_surface_fetch(surface_mask const& m, float x, float y, float z, uint32_t& word,
                                                   uint32_t& bit)
{
    word = 0u;
    bit = 0u;
    if (!m.bits)
        return;
    float const f[3] = {(x - m.lo[0]) * m.inv[0], (y - m.lo[1]) * m.inv[1], (z - m.lo[2]) * m.inv[2]};
    int iv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        if (!(f[a] >= 0.0f && f[a] < float(m.n[a]))) // outside the box, or NaN
        {
            word = 1u;
            return;
        }
        iv[a] = min(int(f[a]), m.n[a] - 1);
    }
    int64_t const v = (int64_t(iv[2]) * m.n[1] + iv[1]) * m.n[0] + iv[0];
    word = m.bits[v >> 5];
    bit = uint32_t(v & 31);
}

__device__ __forceinline__ bool near_surface(surface_mask const& m, float x, float y, float z)
{
    uint32_t w, b;
    near_surface_fetch(m, x, y, z, w, b);
    return (w >> b) & 1u;
}

// Marks the voxels within eps of triangle t: every voxel of the triangle's bounding box grown by eps that also passes
// the separating-plane test along the triangle normal, |n . (c - a)| <= eps + sum_k |n_k| h_k (c: voxel centre, h:
// half voxel). Both tests are necessary conditions for a point of the voxel to lie within eps of the triangle, so the
// mask is conservative; slack covers float rounding.
__global__ void __launch_bounds__(kB) k_mark_surface(float const* __restrict__ v, int32_t const* __restrict__ idx,
                                                     int64_t ntri, surface_mask m, float eps,
                                                     uint32_t* __restrict__ bits)
{
    int64_t const t = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (t >= ntri)
        return;
    float p[3][3];
    for (int c = 0; c < 3; ++c)
        for (int a = 0; a < 3; ++a)
            p[c][a] = v[3 * int64_t(idx[3 * t + c]) + a];
    int i0[3], i1[3];
    for (int a = 0; a < 3; ++a)
    {
        float const lo = fminf(fminf(p[0][a], p[1][a]), p[2][a]) - eps;
        float const hi = fmaxf(fmaxf(p[0][a], p[1][a]), p[2][a]) + eps;
        float const f0 = floorf((lo - m.lo[a]) * m.inv[a]);
        float const f1 = floorf((hi - m.lo[a]) * m.inv[a]);
        if (!(f1 >= 0.0f && f0 < float(m.n[a])))
            return; // entirely outside the box (or NaN coordinates)
        i0[a] = f0 < 0.0f ? 0 : int(f0);
        i1[a] = f1 >= float(m.n[a]) ? m.n[a] - 1 : int(f1);
    }
    double const e1[3] = {double(p[1][0]) - p[0][0], double(p[1][1]) - p[0][1], double(p[1][2]) - p[0][2]};
    double const e2[3] = {double(p[2][0]) - p[0][0], double(p[2][1]) - p[0][1], double(p[2][2]) - p[0][2]};
    double nx = e1[1] * e2[2] - e1[2] * e2[1], ny = e1[2] * e2[0] - e1[0] * e2[2], nz = e1[0] * e2[1] - e1[1] * e2[0];
    double const nl = sqrt(nx * nx + ny * ny + nz * nz);
    bool const plane = nl > 0.0;
    if (plane)
    {
        nx /= nl;
        ny /= nl;
        nz /= nl;
    }
    double const h[3] = {0.5 / m.inv[0], 0.5 / m.inv[1], 0.5 / m.inv[2]};
    double const slack = 1e-5 * (fabs(double(p[0][0])) + fabs(double(p[0][1])) + fabs(double(p[0][2])) + 2.0 * h[0] +
                                 2.0 * h[1] + 2.0 * h[2]);
    double const reach = double(eps) + fabs(nx) * h[0] + fabs(ny) * h[1] + fabs(nz) * h[2] + slack;
    for (int iz = i0[2]; iz <= i1[2]; ++iz)
        for (int iy = i0[1]; iy <= i1[1]; ++iy)
            for (int ix = i0[0]; ix <= i1[0]; ++ix)
            {
                if (plane)
                {
                    double const cx = double(m.lo[0]) + (ix + 0.5) * 2.0 * h[0] - p[0][0];
                    double const cy = double(m.lo[1]) + (iy + 0.5) * 2.0 * h[1] - p[0][1];
                    double const cz = double(m.lo[2]) + (iz + 0.5) * 2.0 * h[2] - p[0][2];
                    if (fabs(nx * cx + ny * cy + nz * cz) > reach)
                        continue;
                }
                int64_t const vi = (int64_t(iz) * m.n[1] + iy) * m.n[0] + ix;
                atomicOr(bits + (vi >> 5), 1u << (vi & 31));
            }
}

__global__ void __launch_bounds__(kB) k_near_surface(float const* __restrict__ pts, int64_t n, surface_mask m,
                                                     int8_t* __restrict__ out)
{
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i < n)
        out[i] = near_surface(m, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]) ? 1 : 0;
}

__device__ __forceinline__ uint32_t hash32(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xFF51AFD7ED558CCDull;
    x ^= x >> 33;
    x *= 0xC4CEB9FE1A85EC53ull;
    x ^= x >> 33;
    return uint32_t(x);
}

// Layer 0 of an axis-separable lattice. When every coordinate of the lattice point start + dx ix + dy iy + dz iz varies
// along one lattice axis at most (each of the three coordinates has at most one nonzero step among dx, dy, dz -- the
// axis-aligned grids, in any axis order), the first layer's input splits by lattice axis: a coordinate's 64 features
// depend on one index only, so
//   z_0(ix, iy, iz) = b_0 + sum_k W_0[k] f_k = T_0[ix] + (T_1[iy] + T_2[iz])
// with T_s[t] the 256-unit sum over the features of the coordinates that vary along lattice axis s, at index t
// (coordinates that vary along none go into T_0, the bias into T_2). The table holds nx + ny + nz rows of 256 FP32
// units (axis s's rows from off[s]); the point's features, the BF16 weights and their exact FP32 products are those
// of the per-point evaluation, only the order of the FP32 sums differs.
struct lattice_map
{
    int32_t dep[3];  // lattice axis along which coordinate a varies, -1: along none
    int64_t off[3];  // first row of lattice axis s: 0, nx, nx + ny
};
constexpr int64_t kMaxLatticeRows = int64_t(1) << 16; // 64 MiB of table at most

__global__ void __launch_bounds__(kH) k_lattice_table(grid_desc const g, feat_params F, lattice_map L,
                                                      __bf16 const* __restrict__ wr0, float const* __restrict__ b0,
                                                      float* __restrict__ T)
{
#pragma clang fp contract(off)
    __shared__ float f[3][2 * kFreq];
    int64_t const r = blockIdx.x;
    int const s = r < L.off[1] ? 0 : (r < L.off[2] ? 1 : 2);
    int64_t const t = r - L.off[s];
    int const n = threadIdx.x;
    bool use[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        use[a] = L.dep[a] == s || (L.dep[a] < 0 && s == 0);
    if (n < 3 * kFreq)
    {
        int const a = n / kFreq, k = n % kFreq;
        if (use[a])
        {
            // the coordinate as query_at forms it: the steps of the other lattice axes are zero for this coordinate,
            // so their terms are exactly zero at every index
            float const ft = float(t);
            float const fi0 = s == 0 ? ft : 0.0f, fi1 = s == 1 ? ft : 0.0f, fi2 = s == 2 ? ft : 0.0f;
            float const p = ((g.start[a] + g.dx[a] * fi0) + g.dy[a] * fi1) + g.dz[a] * fi2;
            float const u = (p - F.c[a]) * F.inv_h[a];
            float const tt = 0.5f * (F.w[k] * u);
            f[a][2 * k] = float(__bf16(__builtin_amdgcn_sinf(tt)));
            f[a][2 * k + 1] = float(__bf16(__builtin_amdgcn_cosf(tt)));
        }
    }
    __syncthreads();
    float acc = s == 2 ? b0[n] : 0.0f;
    for (int a = 0; a < 3; ++a)
        if (use[a])
            for (int q = 0; q < 2 * kFreq; ++q)
                acc = fmaf(f[a][q], float(wr0[int64_t(2 * kFreq * a + q) * kH + n]), acc);
    T[r * kH + n] = acc;
}

// Training samples of the surrogate: point i is drawn from counter-based SplitMix64 streams c = 0..5 of (seed, c), its
// k-th value mix(mix(seed G + c C) + (i + 1) G) (G = 0x9E3779B97F4A7C15, C = 0xD1B54A32D192ED03) turned into a uniform
// double (z >> 11) 2^-53 -- the streams of meshgen.uniform01(seed, c, n). With probability near_frac (stream 0) the
// point lies at a signed normal offset from an area-uniform surface point: the triangle is the first whose cumulative
// doubled area exceeds u1 times the total (cdf: inclusive prefix sums of |(b - a) x (c - a)|), the barycentrics are
// sqrt(u2) and u3, the offset band s u5^power with s = -1 when u4 < 1/2; otherwise it is uniform in the box
// (lo + ext (u1, u2, u3)). Double arithmetic without contraction, rounded to float at the end.
__device__ __forceinline__ uint64_t sm_mix(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ void __launch_bounds__(kB) k_training_samples(float const* __restrict__ v, int32_t const* __restrict__ idx,
                                                         double const* __restrict__ cdf, int64_t ntri, double3 lo,
                                                         double3 ext, int64_t n, double near_frac, double band,
                                                         double power, uint64_t seed, float* __restrict__ out)
{
#pragma clang fp contract(off)
    int64_t const i = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (i >= n)
        return;
    auto uni = [&](uint64_t c) {
        uint64_t const base = sm_mix(seed * 0x9E3779B97F4A7C15ull + c * 0xD1B54A32D192ED03ull);
        return double(sm_mix(base + uint64_t(i + 1) * 0x9E3779B97F4A7C15ull) >> 11) * 0x1p-53;
    };
    double p[3];
    if (ntri > 0 && uni(0) < near_frac)
    {
        double const t = uni(1) * cdf[ntri - 1];
        int64_t a0 = 0, a1 = ntri; // first k with cdf[k] > t
        while (a0 < a1)
        {
            int64_t const m = (a0 + a1) >> 1;
            if (cdf[m] > t)
                a1 = m;
            else
                a0 = m + 1;
        }
        int64_t const k = a0 < ntri ? a0 : ntri - 1;
        double A[3], B[3], C[3];
        for (int d = 0; d < 3; ++d)
        {
            A[d] = v[3 * int64_t(idx[3 * k]) + d];
            B[d] = v[3 * int64_t(idx[3 * k + 1]) + d];
            C[d] = v[3 * int64_t(idx[3 * k + 2]) + d];
        }
        double const r1 = sqrt(uni(2)), r2 = uni(3);
        double const w0 = 1.0 - r1, w1 = r1 * (1.0 - r2), w2 = r1 * r2;
        double const e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double const e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
        double nr[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double const len = sqrt((nr[0] * nr[0] + nr[1] * nr[1]) + nr[2] * nr[2]);
        if (len > 0.0)
            for (int d = 0; d < 3; ++d)
                nr[d] /= len;
        double const off = (band * (uni(4) < 0.5 ? -1.0 : 1.0)) * pow(uni(5), power);
        for (int d = 0; d < 3; ++d)
            p[d] = ((w0 * A[d] + w1 * B[d]) + w2 * C[d]) + off * nr[d];
    }
    else
    {
        p[0] = lo.x + ext.x * uni(1);
        p[1] = lo.y + ext.y * uni(2);
        p[2] = lo.z + ext.z * uni(3);
    }
    for (int d = 0; d < 3; ++d)
        out[3 * i + d] = float(p[d]);
}

template <class Src, bool ROUTE, bool LAT = false>
__global__ void __launch_bounds__(256, kInferBlocks) k_mlp_infer(Src const src, int64_t n, feat_params F, infer_params P,
                                                      float* __restrict__ pred, route_params RP,
                                                      float const* __restrict__ lat)
{
    __shared__ __attribute__((aligned(16))) __bf16 act[kTileP * kLdsStride];
    __shared__ float predw[4][kTileP]; // per wave: the dot over its 64 units
    __shared__ __attribute__((aligned(16))) float par_s[kL * kH + kH]; // biases b_0 .. b_3, then wo
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    int const n0 = wave * 64;
    int64_t const base = int64_t(blockIdx.x) * kTileP;

    // the first MFMA layer's first three weight steps, fetched before anything else so that the prologue hides them
    bf16x8 f0[4], f1[4]; // weight fragments of unit blocks 2w, 2w + 1 for four k steps (see layer)
    {
        int const l = LAT ? 1 : 0;
        int const K = l == 0 ? kIn : kH;
        __bf16 const* wa0 = P.wf[l] + (int64_t(n0 / 32) * (K / 16) * 64 + lane) * 8;
#pragma unroll
        for (int t = 0; t < 3; ++t)
        {
            f0[t] = *reinterpret_cast<bf16x8 const*>(wa0 + 512 * t);
            f1[t] = *reinterpret_cast<bf16x8 const*>(wa0 + int64_t(K / 16) * 512 + 512 * t);
        }
    }
    for (int k = threadIdx.x; k < kL * kH + kH; k += 256)
        par_s[k] = k < kL * kH ? P.b[k / kH][k % kH] : P.wo[k - kL * kH];

    // the routing's near-surface mask word, fetched now and tested after the output layer, so that the MLP hides the
    // load's latency
    uint32_t near_word = 0u, near_bit = 0u;
    if constexpr (ROUTE)
    {
        int64_t const i = base + threadIdx.x;
        if (threadIdx.x < kTileP && i < n)
        {
            float x, y, z;
            query_at(src, i, x, y, z);
            near_surface_fetch(RP.mask, x, y, z, near_word, near_bit);
        }
    }

    int const li = lane & 31;
    int const hl = lane >> 5; // lane half

    // features: the tile's points in kTileJ chunks of 32, a point's 96 pairs in 24 groups of 4 (axis a = g / 8,
    // frequencies 4 (g % 8) .. + 3). Wave w takes groups 6w .. 6w + 5 of every chunk, so the group -- and with it the
    // axis and the frequencies -- is the same across the wave and the frequencies come from scalar registers; lane
    // (li, hl) computes point li of the chunk at frequencies 4 (g % 8) + 2 hl + {0, 1}: 2 pairs, one 8-byte store.
    // Same operations as the training features (feature_pair); points past n get the features of the origin, which
    // nothing reads. (Lattice tables: no features; layer 0 is a sum of three table rows.)
    if constexpr (!LAT)
    {
        float u[kTileJ][3];
#pragma unroll
        for (int c = 0; c < kTileJ; ++c)
        {
            int64_t const i = base + 32 * c + li;
            float x = 0.0f, y = 0.0f, z = 0.0f;
            if (i < n)
                query_at(src, i, x, y, z);
            u[c][0] = (x - F.c[0]) * F.inv_h[0];
            u[c][1] = (y - F.c[1]) * F.inv_h[1];
            u[c][2] = (z - F.c[2]) * F.inv_h[2];
        }
#pragma unroll 1
        for (int gg = 0; gg < 6; ++gg)
        {
            int const g = __builtin_amdgcn_readfirstlane(6 * wave + gg);
            int const a = g >> 3, k0 = 4 * (g & 7);
            float const w0 = hl ? F.w[k0 + 2] : F.w[k0], w1 = hl ? F.w[k0 + 3] : F.w[k0 + 1];
#pragma unroll
            for (int c = 0; c < kTileJ; ++c)
            {
                float const uc = a == 0 ? u[c][0] : (a == 1 ? u[c][1] : u[c][2]);
                float const t0 = 0.5f * (w0 * uc), t1 = 0.5f * (w1 * uc);
                bf16x4 v;
                v[0] = __bf16(__builtin_amdgcn_sinf(t0));
                v[1] = __bf16(__builtin_amdgcn_cosf(t0));
                v[2] = __bf16(__builtin_amdgcn_sinf(t1));
                v[3] = __bf16(__builtin_amdgcn_cosf(t1));
                *reinterpret_cast<bf16x4*>(act + (32 * c + li) * kLdsStride + 2 * (32 * a + k0 + 2 * hl)) = v;
            }
        }
        __syncthreads();
    }

    int const kh = hl * 8;
    f32x16 acc[2][kTileJ]; // C[r][c] of the lane's units: unit r = 8(e/4) + 4(l/32) + e%4
    // layer l's MFMAs: acc = b_l + Wt_l[units of the wave] act^T. The biases (from LDS) go into acc[i][0] and enter
    // every point block's first MFMA as its C operand (no copies). The k loop is unrolled (static buffer indexes; a
    // layer of K = 192 leaves the last 4 of 16 steps). Step s's weight fragments (f[s % 4]) are fetched from L2 three
    // steps ahead -- across layers too: a layer's last three steps fetch the next layer's first three, so no layer
    // starts on an L2 round trip (layers have 12 or 16 steps, multiples of 4, so a step's buffer is the same counted
    // in either layer) -- and its activation operands (b[s & 1], LDS) one step ahead. Scheduling barriers keep the
    // compiler from sinking the fetches back to their uses.
    auto frag = [&](int l) __attribute__((always_inline)) {
        int const K = l == 0 ? kIn : kH;
        return P.wf[l] + (int64_t(n0 / 32) * (K / 16) * 64 + lane) * 8; // unit block 2w; 2w + 1 at + (K / 16) 512
    };
    auto layer = [&](int l) __attribute__((always_inline)) {
        int const K = l == 0 ? kIn : kH;
        int const nst = K / 16;
        __bf16 const* wa0 = frag(l);
        __bf16 const* wa1 = wa0 + int64_t(nst) * 512;
        bool const nxt = l + 1 < kL;
        __bf16 const* na0 = nxt ? frag(l + 1) : wa0;
        __bf16 const* na1 = na0 + int64_t(kH / 16) * 512;
        __bf16 const* brow = act + li * kLdsStride + kh;
        bf16x8 b[2][kTileJ];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                float4 const bq = *reinterpret_cast<float4 const*>(par_s + l * kH + n0 + i * 32 + 8 * q + 4 * hl);
                acc[i][0][4 * q] = bq.x;
                acc[i][0][4 * q + 1] = bq.y;
                acc[i][0][4 * q + 2] = bq.z;
                acc[i][0][4 * q + 3] = bq.w;
            }
#pragma unroll
        for (int j = 0; j < kTileJ; ++j)
            b[0][j] = *reinterpret_cast<bf16x8 const*>(brow + j * 32 * kLdsStride);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int st = 0; st < kH / 16; ++st)
        {
            int const k0 = 16 * st;
            if (k0 >= K)
                break;
            int const buf = st & 1, fb = st % 4;
            if (k0 + 16 < K)
            {
#pragma unroll
                for (int j = 0; j < kTileJ; ++j)
                    b[buf ^ 1][j] = *reinterpret_cast<bf16x8 const*>(brow + j * 32 * kLdsStride + k0 + 16);
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
            {
                bf16x8 const a = i == 0 ? f0[fb] : f1[fb];
                if (st == 0)
                {
#pragma unroll
                    for (int j = kTileJ - 1; j >= 0; --j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[buf][j], acc[i][0], 0, 0, 0);
                }
                else
                {
#pragma unroll
                    for (int j = 0; j < kTileJ; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b[buf][j], acc[i][j], 0, 0, 0);
                }
            }
            int const s3 = st + 3, f3 = s3 % 4;
            if (s3 < nst)
            {
                f0[f3] = *reinterpret_cast<bf16x8 const*>(wa0 + int64_t(s3) * 512);
                f1[f3] = *reinterpret_cast<bf16x8 const*>(wa1 + int64_t(s3) * 512);
            }
            else if (nxt)
            {
                int64_t const t = s3 - nst; // the next layer's step, same buffer
                f0[f3] = *reinterpret_cast<bf16x8 const*>(na0 + t * 512);
                f1[f3] = *reinterpret_cast<bf16x8 const*>(na1 + t * 512);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    // hidden layer l's activations -> LDS (the next layer's B operand)
    auto store_act = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                int const u = n0 + i * 32 + 8 * q + 4 * (lane >> 5); // units u .. u+3
#pragma unroll
                for (int j = 0; j < kTileJ; ++j)
                {
                    bf16x4 hv;
                    hv[0] = __bf16(hsilu(acc[i][j][4 * q]));
                    hv[1] = __bf16(hsilu(acc[i][j][4 * q + 1]));
                    hv[2] = __bf16(hsilu(acc[i][j][4 * q + 2]));
                    hv[3] = __bf16(hsilu(acc[i][j][4 * q + 3]));
                    *reinterpret_cast<bf16x4*>(act + (j * 32 + li) * kLdsStride + u) = hv;
                }
            }
        __syncthreads();
    };
    if constexpr (LAT)
    {
        // layer 0 from the tables: z_0 = T_0[ix] + (T_1[iy] + T_2[iz]) (bias in T_2), activated and stored as
        // store_act does (the accumulators stay free). Wave w fills the LDS rows of points 32w .. 32w + 31, one point
        // per step, lane l taking units 4l .. 4l + 3: every table row is read whole (1 KiB per wave load). Within a
        // wave the 32 points usually share (iy, iz) -- one lattice row -- and then T_1 + T_2 is read once. Points
        // past n take cell (0, 0, 0), which nothing reads.
        src_grid const& gs = src;
        float const* T1 = lat + gs.g.nx * kH;
        float const* T2 = T1 + gs.g.ny * kH;
        int32_t cx = 0, cy = 0, cz = 0; // lane l < 32: the cell of point 32 w + l
        {
            int64_t const i = base + 32 * wave + li;
            if (i < n)
            {
                int64_t ix, iy, iz;
                grid_cell(gs.g, i, ix, iy, iz);
                cx = int32_t(ix);
                cy = int32_t(iy);
                cz = int32_t(iz);
            }
        }
        int const y0 = __builtin_amdgcn_readfirstlane(cy), z0 = __builtin_amdgcn_readfirstlane(cz);
        bool const row = __ballot(cy != y0 || cz != z0) == 0; // uniform across the wave
        __bf16* dst = act + (32 * wave) * kLdsStride + 4 * lane;
        auto put = [&](int p, float4 const a, float4 const bc) __attribute__((always_inline)) {
            bf16x4 hv;
            hv[0] = __bf16(hsilu(a.x + bc.x));
            hv[1] = __bf16(hsilu(a.y + bc.y));
            hv[2] = __bf16(hsilu(a.z + bc.z));
            hv[3] = __bf16(hsilu(a.w + bc.w));
            *reinterpret_cast<bf16x4*>(dst + p * kLdsStride) = hv;
        };
        auto rowsum = [&](int iy, int iz) __attribute__((always_inline)) {
            float4 const b = *reinterpret_cast<float4 const*>(T1 + int64_t(iy) * kH + 4 * lane);
            float4 const c = *reinterpret_cast<float4 const*>(T2 + int64_t(iz) * kH + 4 * lane);
            return float4{b.x + c.x, b.y + c.y, b.z + c.z, b.w + c.w};
        };
        if (row)
        {
            float4 const bc = rowsum(y0, z0);
#pragma unroll 8
            for (int p = 0; p < 32; ++p)
            {
                int const ix = __builtin_amdgcn_readlane(cx, p);
                put(p, *reinterpret_cast<float4 const*>(lat + int64_t(ix) * kH + 4 * lane), bc);
            }
        }
        else
        {
#pragma unroll 4
            for (int p = 0; p < 32; ++p)
            {
                int const ix = __builtin_amdgcn_readlane(cx, p);
                int const iy = __builtin_amdgcn_readlane(cy, p), iz = __builtin_amdgcn_readlane(cz, p);
                put(p, *reinterpret_cast<float4 const*>(lat + int64_t(ix) * kH + 4 * lane), rowsum(iy, iz));
            }
        }
        __syncthreads();
    }
#pragma unroll 1
    for (int l = LAT ? 1 : 0; l + 1 < kL; ++l)
    {
        layer(l);
        __syncthreads(); // every wave has read the layer input
        store_act();
    }
    layer(kL - 1);
    {
        // the output layer fused into the last epilogue: each lane's 32 units of its kTileJ points, activated and
        // rounded to BF16 (the stored activations' values), dotted with wo in FP32; the two lane halves (other units,
        // same points) through a shuffle, the four waves (other units) through LDS
        float part[kTileJ] = {};
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                float4 const wq = *reinterpret_cast<float4 const*>(par_s + kL * kH + n0 + i * 32 + 8 * q + 4 * hl);
#pragma unroll
                for (int j = 0; j < kTileJ; ++j)
                {
                    part[j] = fmaf(float(__bf16(hsilu(acc[i][j][4 * q]))), wq.x, part[j]);
                    part[j] = fmaf(float(__bf16(hsilu(acc[i][j][4 * q + 1]))), wq.y, part[j]);
                    part[j] = fmaf(float(__bf16(hsilu(acc[i][j][4 * q + 2]))), wq.z, part[j]);
                    part[j] = fmaf(float(__bf16(hsilu(acc[i][j][4 * q + 3]))), wq.w, part[j]);
                }
            }
#pragma unroll
        for (int j = 0; j < kTileJ; ++j)
        {
            part[j] += __shfl_xor(part[j], 32);
            if (lane < 32)
                predw[wave][j * 32 + lane] = part[j];
        }
        __syncthreads();
    }
    if (threadIdx.x >= kTileP)
        return;
    int64_t const i = base + threadIdx.x;
    bool const live = i < n;
    float const pv = ((predw[0][threadIdx.x] + predw[1][threadIdx.x]) + (predw[2][threadIdx.x] + predw[3][threadIdx.x])) +
                     *P.bo;
    if (pred && live)
        pred[i] = pv;
    if constexpr (ROUTE)
    {
        bool const near = (near_word >> near_bit) & 1u;
        float const d = fabsf(pv - 0.5f);
        bool const band = live && !near && !(d > RP.delta); // NaN routes too
        bool const audit = live && !near && !band && RP.audit_mod != 0 &&
                           hash32(uint64_t(i) * 0x9E3779B97F4A7C15ull + RP.audit_seed) % RP.audit_mod == 0;
        if (live)
            RP.labels[i] = pv > 0.5f ? 1 : 0;
        bool const take = near || band || audit;
        uint64_t const mtake = __ballot(take);
        if (mtake == 0)
            return;
        uint64_t const mband = __ballot(band);
        uint64_t const mnear = __ballot(near);
        int const leader = __ffsll((unsigned long long)mtake) - 1;
        unsigned long long basei = 0;
        if (lane == leader)
        {
            basei = atomicAdd(RP.counters + 3, (unsigned long long)__popcll(mtake));
            if (mband)
                atomicAdd(RP.counters + 0, (unsigned long long)__popcll(mband));
            if (mnear)
                atomicAdd(RP.counters + 1, (unsigned long long)__popcll(mnear));
            if (mtake & ~(mband | mnear))
                atomicAdd(RP.counters + 2, (unsigned long long)__popcll(mtake & ~(mband | mnear)));
        }
        basei = __shfl(basei, leader);
        if (take)
        {
            int const rank = __popcll(mtake & ((1ull << lane) - 1ull));
            RP.list[basei + rank] = 2 * i + (audit ? 1 : 0);
        }
    }
}

// Applies the exact labels of the routed points: w[e] is the exact GWN of list entry e. Audit entries whose
// surrogate label disagrees are counted in counters[4].
__global__ void __launch_bounds__(kB) k_apply(int64_t const* __restrict__ list, int64_t m, float const* __restrict__ w,
                                              int8_t* __restrict__ labels, unsigned long long* __restrict__ counters)
{
    int64_t const e = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (e >= m)
        return;
    int64_t const code = list[e];
    int64_t const i = code >> 1;
    int8_t const lab = w[e] > 0.5f ? 1 : 0;
    if ((code & 1) && labels[i] != lab)
        atomicAdd(counters + 4, 1ull);
    labels[i] = lab;
}

// gather of routed points: pts[e] = the query of list entry e
template <class Src>
__global__ void __launch_bounds__(kB) k_gather_routed(Src const src, int64_t const* __restrict__ list, int64_t m,
                                                      float* __restrict__ pts)
{
    int64_t const e = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (e >= m)
        return;
    float x, y, z;
    query_at(src, list[e] >> 1, x, y, z);
    pts[3 * e] = x;
    pts[3 * e + 1] = y;
    pts[3 * e + 2] = z;
}
} // namespace

// ---------------------------------------------------------------------------------------------------------------------
// the model object
// ---------------------------------------------------------------------------------------------------------------------
struct mlp_impl
{
    feat_params F{};
    float lo[3]{}, hi[3]{};
    float max_freq = 0.0f;
    float* w = nullptr;  // FP32 master parameters
    float* m1 = nullptr; // Adam moments
    float* m2 = nullptr;
    float* g = nullptr; // gradient
    __bf16* wr[kL] = {};
    __bf16* wt[kL] = {};
    __bf16* wf[kL] = {};
    uint32_t* mask_bits = nullptr;
    surface_mask mask{};
    float mask_eps = 0.0f;
    int64_t step = 0;
    std::mutex mu;

    mlp_impl() = default;
    mlp_impl(mlp_impl const&) = delete;
    ~mlp_impl()
    {
        for (float* p : {w, m1, m2, g})
            if (p)
                (void)hipFree(p);
        if (mask_bits)
            (void)hipFree(mask_bits);
        for (int l = 0; l < kL; ++l)
        {
            if (wr[l])
                (void)hipFree(wr[l]);
            if (wt[l])
                (void)hipFree(wt[l]);
            if (wf[l])
                (void)hipFree(wf[l]);
        }
    }

    int alloc()
    {
        size_t const P = size_t(kLay.total) * sizeof(float);
        AGWN_HIP(hipMalloc(&w, P));
        AGWN_HIP(hipMalloc(&m1, P));
        AGWN_HIP(hipMalloc(&m2, P));
        AGWN_HIP(hipMalloc(&g, P));
        AGWN_HIP(hipMemset(m1, 0, P));
        AGWN_HIP(hipMemset(m2, 0, P));
        AGWN_HIP(hipMemset(g, 0, P));
        for (int l = 0; l < kL; ++l)
        {
            AGWN_HIP(hipMalloc(&wr[l], size_t(kInDim[l] * kH) * sizeof(__bf16)));
            AGWN_HIP(hipMalloc(&wt[l], size_t(kInDim[l] * kH) * sizeof(__bf16)));
            AGWN_HIP(hipMalloc(&wf[l], size_t(kInDim[l] * kH) * sizeof(__bf16)));
        }
        return ok;
    }

    wcopies copies() const
    {
        wcopies C{};
        for (int l = 0; l < kL; ++l)
        {
            C.wr[l] = wr[l];
            C.wt[l] = wt[l];
            C.wf[l] = wf[l];
        }
        return C;
    }

    int cast(hipStream_t s)
    {
        k_cast<<<dim3(unsigned((kLay.total + kB - 1) / kB)), dim3(kB), 0, s>>>(w, kLay.total, copies());
        AGWN_HIP(hipGetLastError());
        return ok;
    }

    infer_params iparams() const
    {
        infer_params P{};
        for (int l = 0; l < kL; ++l)
        {
            P.wf[l] = wf[l];
            P.b[l] = w + kLay.b[l];
        }
        P.wo = w + kLay.wo;
        P.bo = w + kLay.bo;
        return P;
    }
};
} // namespace agwn

struct agwn_mlp
{
    agwn::mlp_impl impl;
};

namespace agwn
{
namespace
{
uint64_t splitmix(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

feat_params make_feats(float const* lo, float const* hi, float max_freq)
{
    feat_params F{};
    for (int a = 0; a < 3; ++a)
    {
        F.c[a] = 0.5f * (lo[a] + hi[a]);
        F.inv_h[a] = 1.0f / (0.5f * (hi[a] - lo[a]));
    }
    for (int k = 0; k < kFreq; ++k)
        F.w[k] = float(std::exp2(std::log2(double(max_freq)) * double(k) / double(kFreq - 1)));
    return F;
}

// Scratch of one forward/backward pass over `rows` samples.
struct train_ws
{
    int64_t rows = 0, parts = 0, kchunk = 0;
    dbuf X0, X0T;       // features [rows x 192], [192 x rows]
    dbuf H[kL], HT[kL]; // BF16 activations and their transposes
    dbuf Z[kL];         // FP32 pre-activations [rows x 256]
    dbuf D[2], DT[2];   // BF16 backward activations (ping-pong over layers)
    dbuf part, pwo, psc, perm, loss;
    int alloc(int64_t r, hipStream_t s)
    {
        rows = r;
        // split the batch reduction of dW so that its launch has >= 256 workgroups
        kchunk = std::max<int64_t>(256, ((rows / 64) + 15) / 16 * 16);
        parts = (rows + kchunk - 1) / kchunk;
        size_t const R = size_t(rows);
        AGWN_TRY(X0.alloc(R * kIn * sizeof(__bf16), s));
        AGWN_TRY(X0T.alloc(R * kIn * sizeof(__bf16), s));
        for (int l = 0; l < kL; ++l)
        {
            AGWN_TRY(H[l].alloc(R * kH * sizeof(__bf16), s));
            AGWN_TRY(HT[l].alloc(R * kH * sizeof(__bf16), s));
            AGWN_TRY(Z[l].alloc(R * kH * sizeof(float), s));
        }
        for (int i = 0; i < 2; ++i)
        {
            AGWN_TRY(D[i].alloc(R * kH * sizeof(__bf16), s));
            AGWN_TRY(DT[i].alloc(R * kH * sizeof(__bf16), s));
        }
        AGWN_TRY(part.alloc(size_t(parts) * kH * kH * sizeof(float), s));
        AGWN_TRY(pwo.alloc(size_t(rows / kOblRows) * kH * sizeof(float), s));
        AGWN_TRY(psc.alloc(size_t(rows / kOblRows) * 2 * sizeof(float), s));
        AGWN_TRY(perm.alloc(R * sizeof(int64_t), s));
        AGWN_TRY(loss.alloc(sizeof(double), s));
        return ok;
    }
};

// Forward (and, with y, backward) pass of the rows whose features are in ws.X0. The gradient lands in M.g, the
// batch's squared-error sum is added to *ws.loss.
int pass(mlp_impl& M, train_ws& ws, int64_t nvalid, float const* y, int64_t const* perm, float* pred, hipStream_t s)
{
    int64_t const rows = ws.rows;
    dim3 const blk(256);
    bool const bwd = y != nullptr;
    for (int l = 0; l < kL; ++l)
    {
        int64_t const K = kInDim[l];
        __bf16 const* x = l == 0 ? ws.X0.as<__bf16>() : ws.H[l - 1].as<__bf16>();
        epi_args E{};
        E.bias = M.w + kLay.b[l];
        E.zout = ws.Z[l].as<float>();
        E.h = ws.H[l].as<__bf16>();
        E.ht = ws.HT[l].as<__bf16>();
        E.rows = rows;
        k_gemm<EPI_FWD><<<dim3(unsigned(rows / 128), kH / 128, 1), blk, 0, s>>>(M.wt[l], K, x, K, kH, rows, K, K, E);
    }
    int const top = (kL - 1) & 1;
    k_out_loss<<<dim3(unsigned(rows / kOblRows)), blk, 0, s>>>(
        ws.H[kL - 1].as<__bf16>(), ws.Z[kL - 1].as<float>(), M.w + kLay.wo, M.w + kLay.bo, y, perm, nvalid, rows, pred,
        bwd ? ws.D[top].as<__bf16>() : nullptr, bwd ? ws.DT[top].as<__bf16>() : nullptr, ws.pwo.as<float>(),
        ws.psc.as<float>());
    AGWN_HIP(hipGetLastError());
    if (!bwd)
        return ok;
    k_out_final<<<dim3(1), blk, 0, s>>>(ws.pwo.as<float>(), ws.psc.as<float>(), rows / kOblRows, M.g + kLay.wo, M.g + kLay.bo,
                                        ws.loss.as<double>());
    for (int l = kL - 1; l >= 0; --l)
    {
        int64_t const K = kInDim[l];
        __bf16 const* xt = l == 0 ? ws.X0T.as<__bf16>() : ws.HT[l - 1].as<__bf16>();
        __bf16 const* dt = ws.DT[l & 1].as<__bf16>();
        // dW_l[k][n] = sum_m XT_l[k][m] DT_l[n][m], split over the batch, partials summed in a fixed order
        epi_args E{};
        E.part = ws.part.as<float>();
        k_gemm<EPI_PART><<<dim3(kH / 128, unsigned((K + 127) / 128), unsigned(ws.parts)), blk, 0, s>>>(
            xt, rows, dt, rows, K, kH, rows, ws.kchunk, E);
        k_sum_parts<<<dim3(unsigned((K * kH + kB - 1) / kB)), blk, 0, s>>>(ws.part.as<float>(), ws.parts, K * kH,
                                                                           M.g + kLay.w[l]);
        k_rowsum<<<dim3(kH), blk, 0, s>>>(dt, rows, M.g + kLay.b[l]);
        if (l > 0)
        {
            // D_{l-1}[m][k] = (sum_n Wr_l[k][n] D_l[m][n]) hsilu'(Z_{l-1}[m][k])
            epi_args E2{};
            E2.zin = ws.Z[l - 1].as<float>();
            E2.h = ws.D[(l - 1) & 1].as<__bf16>();
            E2.ht = ws.DT[(l - 1) & 1].as<__bf16>();
            E2.rows = rows;
            k_gemm<EPI_DX><<<dim3(unsigned(rows / 128), unsigned((K + 127) / 128), 1), blk, 0, s>>>(
                M.wr[l], kH, ws.D[l & 1].as<__bf16>(), kH, K, rows, kH, kH, E2);
        }
    }
    AGWN_HIP(hipGetLastError());
    return ok;
}

int64_t gcd64(int64_t a, int64_t b)
{
    while (b)
    {
        int64_t const t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int mlp_create(float const* lo, float const* hi, float max_freq, uint64_t seed, agwn_mlp** out)
{
    AGWN_CHECK(out && lo && hi, "null argument");
    *out = nullptr;
    AGWN_CHECK(std::isfinite(max_freq) && max_freq >= 1.0f && max_freq <= 4096.0f,
               "max_freq must be in [1, 4096] (got %g)", double(max_freq));
    for (int a = 0; a < 3; ++a)
        AGWN_CHECK(std::isfinite(lo[a]) && std::isfinite(hi[a]) && hi[a] > lo[a], "empty or non-finite box on axis %d",
                   a);
    AGWN_TRY(ensure_device());
    auto h = std::make_unique<agwn_mlp>();
    mlp_impl& M = h->impl;
    M.F = make_feats(lo, hi, max_freq);
    M.max_freq = max_freq;
    for (int a = 0; a < 3; ++a)
    {
        M.lo[a] = lo[a];
        M.hi[a] = hi[a];
    }
    AGWN_TRY(M.alloc());
    // He-uniform hidden layers, small output layer, zero biases, bo = 0.5 (the undecided value)
    std::vector<float> w(size_t(kLay.total), 0.0f);
    uint64_t st = seed ^ 0x6A09E667F3BCC908ull;
    auto uni = [&] { return float(double(splitmix(st) >> 11) * 0x1p-53) * 2.0f - 1.0f; };
    for (int l = 0; l < kL; ++l)
    {
        float const a = std::sqrt(6.0f / float(kInDim[l]));
        for (int64_t i = 0; i < kInDim[l] * kH; ++i)
            w[size_t(kLay.w[l] + i)] = a * uni();
    }
    float const ao = 0.1f * std::sqrt(6.0f / float(kH));
    for (int j = 0; j < kH; ++j)
        w[size_t(kLay.wo + j)] = ao * uni();
    w[size_t(kLay.bo)] = 0.5f;
    AGWN_HIP(hipMemcpy(M.w, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
    AGWN_TRY(M.cast(nullptr));
    AGWN_HIP(hipDeviceSynchronize());
    *out = h.release();
    return ok;
}

int mlp_set_params(agwn_mlp* h, float const* p)
{
    AGWN_CHECK(h && p, "null argument");
    mlp_impl& M = h->impl;
    std::lock_guard<std::mutex> lk(M.mu);
    size_t const P = size_t(kLay.total) * sizeof(float);
    AGWN_HIP(hipMemcpy(M.w, p, P, hipMemcpyHostToDevice));
    AGWN_HIP(hipMemset(M.m1, 0, P));
    AGWN_HIP(hipMemset(M.m2, 0, P));
    M.step = 0;
    AGWN_TRY(M.cast(nullptr));
    AGWN_HIP(hipDeviceSynchronize());
    return ok;
}

int mlp_get_params(agwn_mlp const* h, float* p)
{
    AGWN_CHECK(h && p, "null argument");
    AGWN_HIP(hipMemcpy(p, h->impl.w, size_t(kLay.total) * sizeof(float), hipMemcpyDeviceToHost));
    return ok;
}

int mlp_train(agwn_mlp* h, float const* pts, float const* y, int64_t n, int batch, int epochs, float lr, uint64_t seed,
              float* loss_host, hipStream_t s)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    AGWN_CHECK(pts && y && n > 0, "empty training set");
    AGWN_CHECK(batch >= 128 && batch % 128 == 0 && batch <= (1 << 22),
               "batch must be a multiple of 128 in [128, 2^22] (got %d)", batch);
    AGWN_CHECK(epochs >= 0, "negative epochs");
    AGWN_CHECK(std::isfinite(lr) && lr > 0.0f, "learning rate must be positive");
    mlp_impl& M = h->impl;
    std::lock_guard<std::mutex> lk(M.mu);
    train_ws ws;
    AGWN_TRY(ws.alloc(batch, s));
    int64_t const B = batch;
    int64_t const steps_per_epoch = (n + B - 1) / B;
    int64_t const total = steps_per_epoch * epochs;
    float const b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    uint64_t st = seed ^ 0xBB67AE8584CAA73Bull;
    wcopies const C = M.copies();
    int64_t it = 0;
    for (int e = 0; e < epochs; ++e)
    {
        int64_t stride = int64_t(splitmix(st) % uint64_t(n)) | 1;
        while (gcd64(stride, n) != 1)
            stride += 2;
        int64_t const offset = int64_t(splitmix(st) % uint64_t(n));
        AGWN_HIP(hipMemsetAsync(ws.loss.p, 0, sizeof(double), s));
        for (int64_t b = 0; b < steps_per_epoch; ++b, ++it)
        {
            k_perm<<<dim3(unsigned((B + kB - 1) / kB)), dim3(kB), 0, s>>>(ws.perm.as<int64_t>(), B, b * B, offset,
                                                                         stride, n);
            k_features<<<dim3(unsigned((B * (kIn / 2) + kB - 1) / kB)), dim3(kB), 0, s>>>(
                pts, ws.perm.as<int64_t>(), B, B, M.F, ws.X0.as<__bf16>(), ws.X0T.as<__bf16>());
            AGWN_HIP(hipGetLastError());
            AGWN_TRY(pass(M, ws, B, y, ws.perm.as<int64_t>(), nullptr, s));
            ++M.step;
            // cosine decay from lr to lr / 20 over the run
            double const frac = total > 1 ? double(it) / double(total - 1) : 0.0;
            float const lr_t = float(double(lr) * (0.05 + 0.475 * (1.0 + std::cos(3.141592653589793 * frac))));
            float const c1 = float(1.0 / (1.0 - std::pow(double(b1), double(M.step))));
            float const c2 = float(1.0 / (1.0 - std::pow(double(b2), double(M.step))));
            k_adam<<<dim3(unsigned((kLay.total + kB - 1) / kB)), dim3(kB), 0, s>>>(M.w, M.m1, M.m2, M.g, kLay.total,
                                                                                   lr_t, b1, b2, eps, c1, c2, C);
            AGWN_HIP(hipGetLastError());
        }
        if (loss_host)
        {
            double L = 0.0;
            AGWN_HIP(hipMemcpyAsync(&L, ws.loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
            AGWN_HIP(hipStreamSynchronize(s));
            loss_host[e] = float(L / double(steps_per_epoch * B));
        }
    }
    AGWN_HIP(hipStreamSynchronize(s));
    return ok;
}

// One forward + backward pass over host samples (no update): gradient of the mean squared error, predictions, loss.
int mlp_grad(agwn_mlp* h, float const* pts, float const* y, int64_t n, float* grad, float* pred, float* loss)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    AGWN_CHECK(pts && y && n > 0 && n <= (1 << 20), "n must be in [1, 2^20]");
    AGWN_CHECK(grad || pred || loss, "no output requested");
    mlp_impl& M = h->impl;
    std::lock_guard<std::mutex> lk(M.mu);
    hipStream_t s = nullptr;
    AGWN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct sguard
    {
        hipStream_t s;
        ~sguard() { (void)hipStreamDestroy(s); }
    } sg{s};
    int64_t const rows = (n + 127) / 128 * 128;
    {
        train_ws ws;
        AGWN_TRY(ws.alloc(rows, s));
        dbuf dp, dy, dpred;
        AGWN_TRY(dp.alloc(size_t(n) * 3 * sizeof(float), s));
        AGWN_TRY(dy.alloc(size_t(n) * sizeof(float), s));
        AGWN_TRY(dpred.alloc(size_t(n) * sizeof(float), s));
        AGWN_HIP(hipMemcpyAsync(dp.p, pts, size_t(n) * 3 * sizeof(float), hipMemcpyHostToDevice, s));
        AGWN_HIP(hipMemcpyAsync(dy.p, y, size_t(n) * sizeof(float), hipMemcpyHostToDevice, s));
        AGWN_HIP(hipMemsetAsync(ws.loss.p, 0, sizeof(double), s));
        k_features<<<dim3(unsigned((rows * (kIn / 2) + kB - 1) / kB)), dim3(kB), 0, s>>>(
            dp.as<float>(), nullptr, n, rows, M.F, ws.X0.as<__bf16>(), ws.X0T.as<__bf16>());
        AGWN_HIP(hipGetLastError());
        AGWN_TRY(pass(M, ws, n, dy.as<float>(), nullptr, dpred.as<float>(), s));
        if (grad)
            AGWN_HIP(hipMemcpyAsync(grad, M.g, size_t(kLay.total) * sizeof(float), hipMemcpyDeviceToHost, s));
        if (pred)
            AGWN_HIP(hipMemcpyAsync(pred, dpred.p, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, s));
        double L = 0.0;
        AGWN_HIP(hipMemcpyAsync(&L, ws.loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
        AGWN_HIP(hipStreamSynchronize(s));
        if (loss)
            *loss = float(L / double(n));
    }
    AGWN_HIP(hipStreamSynchronize(s));
    return ok;
}

// Whether layer 0 of the lattice splits by lattice axis (see k_lattice_table), and how: each coordinate has at most one
// nonzero step among dx, dy, dz, and the table's nx + ny + nz rows stay within kMaxLatticeRows.
bool lattice_separable(grid_desc const& g, lattice_map& L)
{
    if (g.nx + g.ny + g.nz > kMaxLatticeRows)
        return false;
    for (int a = 0; a < 3; ++a)
    {
        float const d[3] = {g.dx[a], g.dy[a], g.dz[a]};
        L.dep[a] = -1;
        for (int k = 0; k < 3; ++k)
            if (d[k] != 0.0f)
            {
                if (L.dep[a] >= 0)
                    return false;
                L.dep[a] = k;
            }
    }
    L.off[0] = 0;
    L.off[1] = g.nx;
    L.off[2] = g.nx + g.ny;
    return true;
}

template <class Src, bool ROUTE>
int infer_launch(mlp_impl const& M, Src const& src, int64_t n, float* pred, route_params const& RP, hipStream_t s,
                 float* lat = nullptr)
{
    if (n == 0)
        return ok;
    int64_t const blocks = (n + kTileP - 1) / kTileP;
    AGWN_CHECK(blocks < (int64_t(1) << 31), "too many points (%lld)", (long long)n);
    if constexpr (std::is_same_v<Src, src_grid>)
    {
        lattice_map L{};
        if (lat && lattice_separable(src.g, L))
        {
            int64_t const rows = src.g.nx + src.g.ny + src.g.nz;
            k_lattice_table<<<dim3(unsigned(rows)), dim3(kH), 0, s>>>(src.g, M.F, L, M.wr[0], M.w + kLay.b[0], lat);
            AGWN_HIP(hipGetLastError());
            k_mlp_infer<Src, ROUTE, true><<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(src, n, M.F, M.iparams(), pred,
                                                                                      RP, lat);
            AGWN_HIP(hipGetLastError());
            return ok;
        }
    }
    k_mlp_infer<Src, ROUTE><<<dim3(unsigned(blocks)), dim3(256), 0, s>>>(src, n, M.F, M.iparams(), pred, RP, nullptr);
    AGWN_HIP(hipGetLastError());
    return ok;
}
} // namespace

int64_t mlp_num_params() { return kLay.total; }

// Builds the near-surface mask of the network's box: voxels of edge `voxel` (<= 0: eps / 2) that come within eps of a
// triangle of the mesh (host arrays).
int mlp_set_surface(agwn_mlp* h, float const* verts, int64_t nv, int32_t const* idx, int64_t nidx, float eps,
                    float voxel)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    AGWN_CHECK(nv >= 0 && nidx >= 0 && nidx % 3 == 0, "bad mesh sizes (nv=%lld, nidx=%lld)", (long long)nv,
               (long long)nidx);
    AGWN_CHECK(nidx == 0 || (verts && idx), "null mesh arrays");
    AGWN_CHECK(std::isfinite(eps) && eps >= 0.0f, "eps must be finite and >= 0");
    for (int64_t i = 0; i < nidx; ++i)
        AGWN_CHECK(idx[i] >= 0 && idx[i] < nv, "index %lld = %d out of range [0, %lld)", (long long)i, idx[i],
                   (long long)nv);
    if (!(voxel > 0.0f))
        voxel = eps > 0.0f ? 0.5f * eps : 0.0f;
    mlp_impl& M = h->impl;
    surface_mask m{};
    int64_t nvox = 1;
    for (int a = 0; a < 3; ++a)
    {
        double const ext = double(M.hi[a]) - double(M.lo[a]);
        double const cells = voxel > 0.0f ? std::ceil(ext / double(voxel)) : 1.0;
        AGWN_CHECK(cells <= double(1 << 20), "voxel %g too small for the box", double(voxel));
        m.n[a] = std::max(1, int(cells));
        m.lo[a] = M.lo[a];
        m.inv[a] = float(double(m.n[a]) / ext);
        nvox *= m.n[a];
    }
    AGWN_CHECK(nvox <= (int64_t(1) << 33), "mask of %lld voxels is too large", (long long)nvox);
    std::lock_guard<std::mutex> lk(M.mu);
    hipStream_t s = nullptr;
    AGWN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct sguard
    {
        hipStream_t s;
        ~sguard() { (void)hipStreamDestroy(s); }
    } sg{s};
    size_t const words = size_t((nvox + 31) / 32);
    uint32_t* bits = nullptr;
    AGWN_HIP(hipMalloc(&bits, words * sizeof(uint32_t)));
    struct bguard
    {
        uint32_t* p;
        ~bguard()
        {
            if (p)
                (void)hipFree(p);
        }
    } bg{bits};
    AGWN_HIP(hipMemsetAsync(bits, 0, words * sizeof(uint32_t), s));
    int64_t const ntri = nidx / 3;
    if (ntri > 0)
    {
        dbuf dv, di;
        AGWN_TRY(dv.alloc(size_t(nv) * 3 * sizeof(float), s));
        AGWN_TRY(di.alloc(size_t(nidx) * sizeof(int32_t), s));
        AGWN_HIP(hipMemcpyAsync(dv.p, verts, size_t(nv) * 3 * sizeof(float), hipMemcpyHostToDevice, s));
        AGWN_HIP(hipMemcpyAsync(di.p, idx, size_t(nidx) * sizeof(int32_t), hipMemcpyHostToDevice, s));
        k_mark_surface<<<dim3(unsigned((ntri + kB - 1) / kB)), dim3(kB), 0, s>>>(dv.as<float>(), di.as<int32_t>(), ntri,
                                                                                m, eps, bits);
        AGWN_HIP(hipGetLastError());
    }
    AGWN_HIP(hipStreamSynchronize(s));
    if (M.mask_bits)
        AGWN_HIP(hipFree(M.mask_bits));
    M.mask_bits = bits;
    bg.p = nullptr;
    m.bits = bits;
    M.mask = m;
    M.mask_eps = eps;
    return ok;
}

int training_samples_dev(float const* verts, int32_t const* idx, double const* cdf, int64_t ntri, float const* lo,
                         float const* hi, int64_t n, double near_frac, double band, double power, uint64_t seed,
                         float* out, hipStream_t s)
{
    AGWN_CHECK(n >= 0 && ntri >= 0, "negative count");
    AGWN_CHECK(lo && hi, "null box");
    AGWN_CHECK(n == 0 || out, "null output");
    AGWN_CHECK(ntri == 0 || (verts && idx && cdf), "null mesh arrays");
    AGWN_CHECK(near_frac >= 0.0 && near_frac <= 1.0 && band >= 0.0 && power > 0.0 && std::isfinite(band) &&
                   std::isfinite(power),
               "bad sampling parameters (near_frac %g, band %g, power %g)", near_frac, band, power);
    for (int a = 0; a < 3; ++a)
        AGWN_CHECK(std::isfinite(lo[a]) && std::isfinite(hi[a]) && lo[a] <= hi[a], "bad box on axis %d", a);
    if (n == 0)
        return ok;
    double3 const L{double(lo[0]), double(lo[1]), double(lo[2])};
    double3 const E{double(hi[0]) - double(lo[0]), double(hi[1]) - double(lo[1]), double(hi[2]) - double(lo[2])};
    k_training_samples<<<dim3(unsigned((n + kB - 1) / kB)), dim3(kB), 0, s>>>(verts, idx, cdf, ntri, L, E, n, near_frac,
                                                                              band, power, seed, out);
    AGWN_HIP(hipGetLastError());
    return ok;
}

int mlp_near_surface_dev(agwn_mlp const* h, float const* pts, int64_t n, int8_t* out, hipStream_t s)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    AGWN_CHECK(n >= 0 && (n == 0 || (pts && out)), "bad arrays");
    AGWN_CHECK(h->impl.mask.bits != nullptr, "the surface mask was not set");
    if (n == 0)
        return ok;
    k_near_surface<<<dim3(unsigned((n + kB - 1) / kB)), dim3(kB), 0, s>>>(pts, n, h->impl.mask, out);
    AGWN_HIP(hipGetLastError());
    return ok;
}

int mlp_predict_dev(agwn_mlp const* h, float const* pts, int64_t n, float* out, hipStream_t s)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    AGWN_CHECK(n >= 0, "negative count");
    AGWN_CHECK(n == 0 || (pts && out), "null array");
    return infer_launch<src_array<float>, false>(h->impl, src_array<float>{pts}, n, out, route_params{}, s);
}

int64_t mlp_lattice_table_floats(grid_desc const& g)
{
    lattice_map L{};
    return lattice_separable(g, L) ? (g.nx + g.ny + g.nz) * kH : 0;
}

int mlp_predict_grid_dev(agwn_mlp const* h, grid_desc const& g, float* out, hipStream_t s)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    AGWN_CHECK(out != nullptr, "null output");
    int64_t const n = g.nx * g.ny * g.nz;
    int64_t const tf = mlp_lattice_table_floats(g);
    dbuf lat;
    if (tf > 0)
        AGWN_TRY(lat.alloc(size_t(tf) * sizeof(float), s));
    return infer_launch<src_grid, false>(h->impl, src_grid{g}, n, out, route_params{}, s, lat.as<float>());
}

template <class Src>
int mlp_route(agwn_mlp const* h, Src const& src, int64_t n, float delta, uint32_t audit_mod, uint32_t audit_seed,
              int8_t* labels, int64_t* list, unsigned long long* counters, hipStream_t s, float* lat)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    route_params RP{delta, audit_mod, audit_seed, labels, list, counters, h->impl.mask};
    return infer_launch<Src, true>(h->impl, src, n, nullptr, RP, s, lat);
}
template int mlp_route<src_array<float>>(agwn_mlp const*, src_array<float> const&, int64_t, float, uint32_t,
                                         uint32_t, int8_t*, int64_t*, unsigned long long*, hipStream_t, float*);
template int mlp_route<src_grid>(agwn_mlp const*, src_grid const&, int64_t, float, uint32_t, uint32_t, int8_t*,
                                 int64_t*, unsigned long long*, hipStream_t, float*);

int route_apply(int64_t const* list, int64_t m, float const* w, int8_t* labels, unsigned long long* counters,
                hipStream_t s)
{
    if (m == 0)
        return ok;
    k_apply<<<dim3(unsigned((m + kB - 1) / kB)), dim3(kB), 0, s>>>(list, m, w, labels, counters);
    AGWN_HIP(hipGetLastError());
    return ok;
}

// Routed-list order: spatially coherent runs make the exact kernels' waves coherent (the integer term's traversals
// and the far-field decision's exactly summed loops). Key = 30-bit Morton code of the query quantised to 1024 cells
// per axis of the network's box (points outside clamp to the border cells), above the entry's own bits.
__device__ __forceinline__ uint64_t spread3_10(uint32_t v)
{
    uint64_t x = v & 0x3ffu;
    x = (x | (x << 16)) & 0x030000FFull;
    x = (x | (x << 8)) & 0x0300F00Full;
    x = (x | (x << 4)) & 0x030C30C3ull;
    x = (x | (x << 2)) & 0x09249249ull;
    return x;
}
struct key_box
{
    float lo[3], scale[3]; // cell = (p - lo) * scale, clamped to [0, 1023]
};
// keys[e] = (Morton code of the query of entry e of the index-ordered list) << pbits | e
template <class Src>
__global__ void __launch_bounds__(kB) k_route_keys(Src const src, uint64_t const* __restrict__ list, int64_t m,
                                                   key_box K, int pbits, uint64_t* __restrict__ keys)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k >= m)
        return;
    float p[3];
    query_at(src, int64_t(list[k] >> 1), p[0], p[1], p[2]);
    uint64_t code = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
    {
        float const c = fminf(fmaxf((p[a] - K.lo[a]) * K.scale[a], 0.0f), 1023.0f); // NaN -> 0
        code |= spread3_10(uint32_t(c)) << a;
    }
    keys[k] = (code << pbits) | uint64_t(k);
}
__global__ void __launch_bounds__(kB) k_strip_keys(uint64_t* __restrict__ keys, int64_t m, uint64_t mask)
{
    int64_t const k = int64_t(blockIdx.x) * kB + threadIdx.x;
    if (k < m)
        keys[k] &= mask;
}

template <class Src>
int route_order(agwn_mlp const* h, Src const& src, int64_t n, int64_t* list, int64_t m, uint64_t* tmp,
                uint64_t* keys, uint64_t* keys2, uint32_t* sort_ws, hipStream_t s, int64_t const** by_index,
                uint64_t const** perm)
{
    AGWN_CHECK(h != nullptr, "mlp handle is null");
    *by_index = list;
    if (perm)
        *perm = nullptr;
    if (m == 0)
        return ok;
    int ebits = 1;
    while (ebits < 63 && (uint64_t(2 * n) >> ebits) != 0)
        ++ebits;
    uint64_t* sorted = nullptr;
    AGWN_TRY(radix_sort_u64_ws(reinterpret_cast<uint64_t*>(list), tmp, m, ebits, sort_ws, s, &sorted));
    *by_index = reinterpret_cast<int64_t const*>(sorted);
    int pbits = 1;
    while (pbits < 63 && (uint64_t(m - 1) >> pbits) != 0)
        ++pbits;
    if (!perm || pbits + 30 > 64) // index order only
        return ok;
    mlp_impl const& M = h->impl;
    key_box K{};
    for (int a = 0; a < 3; ++a)
    {
        double const ext = double(M.hi[a]) - double(M.lo[a]);
        K.lo[a] = M.lo[a];
        K.scale[a] = ext > 0.0 ? float(1024.0 / ext) : 0.0f;
    }
    unsigned const g = unsigned((m + kB - 1) / kB);
    k_route_keys<Src><<<dim3(g), dim3(kB), 0, s>>>(src, sorted, m, K, pbits, keys);
    AGWN_HIP(hipGetLastError());
    uint64_t* ks = nullptr;
    AGWN_TRY(radix_sort_u64_ws(keys, keys2, m, pbits + 30, sort_ws, s, &ks));
    k_strip_keys<<<dim3(g), dim3(kB), 0, s>>>(ks, m, (uint64_t(1) << pbits) - 1);
    AGWN_HIP(hipGetLastError());
    *perm = ks;
    return ok;
}
template int route_order<src_array<float>>(agwn_mlp const*, src_array<float> const&, int64_t, int64_t*, int64_t,
                                           uint64_t*, uint64_t*, uint64_t*, uint32_t*, hipStream_t, int64_t const**,
                                           uint64_t const**);
template int route_order<src_grid>(agwn_mlp const*, src_grid const&, int64_t, int64_t*, int64_t, uint64_t*, uint64_t*,
                                   uint64_t*, uint32_t*, hipStream_t, int64_t const**, uint64_t const**);

template <class Src>
int gather_routed(Src const& src, int64_t const* list, int64_t m, float* pts, hipStream_t s)
{
    if (m == 0)
        return ok;
    k_gather_routed<Src><<<dim3(unsigned((m + kB - 1) / kB)), dim3(kB), 0, s>>>(src, list, m, pts);
    AGWN_HIP(hipGetLastError());
    return ok;
}
template int gather_routed<src_array<float>>(src_array<float> const&, int64_t const*, int64_t, float*, hipStream_t);
template int gather_routed<src_grid>(src_grid const&, int64_t const*, int64_t, float*, hipStream_t);
} // namespace agwn

// ---------------------------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------------------------
namespace
{
template <class F>
int guarded_mlp(F&& f)
{
    agwn::clear_error();
    try
    {
        return f();
    }
    catch (std::bad_alloc const&)
    {
        agwn::set_error("host allocation failed");
        return agwn::err_alloc;
    }
    catch (...)
    {
        agwn::set_error("unexpected C++ exception");
        return agwn::err_hip;
    }
}
} // namespace

extern "C" {
int64_t agwn_mlp_num_params(void) { return agwn::mlp_num_params(); }
int agwn_mlp_create(float const* lo, float const* hi, float max_freq, uint64_t seed, agwn_mlp** out)
{
    return guarded_mlp([&] { return agwn::mlp_create(lo, hi, max_freq, seed, out); });
}
int agwn_mlp_destroy(agwn_mlp* m)
{
    agwn::clear_error();
    delete m;
    return agwn::ok;
}
int agwn_mlp_info(agwn_mlp const* m, float* lo, float* hi, float* max_freq)
{
    return guarded_mlp([&]() -> int {
        AGWN_CHECK(m != nullptr, "mlp handle is null");
        for (int a = 0; a < 3; ++a)
        {
            if (lo)
                lo[a] = m->impl.lo[a];
            if (hi)
                hi[a] = m->impl.hi[a];
        }
        if (max_freq)
            *max_freq = m->impl.max_freq;
        return agwn::ok;
    });
}
int agwn_mlp_get_params(agwn_mlp const* m, float* out)
{
    return guarded_mlp([&] { return agwn::mlp_get_params(m, out); });
}
int agwn_mlp_set_params(agwn_mlp* m, float const* in)
{
    return guarded_mlp([&] { return agwn::mlp_set_params(m, in); });
}
int agwn_mlp_train_dev(agwn_mlp* m, float const* pts, float const* target, int64_t n, int batch, int epochs, float lr,
                       uint64_t seed, float* loss_per_epoch, agwn_stream s)
{
    return guarded_mlp(
        [&] { return agwn::mlp_train(m, pts, target, n, batch, epochs, lr, seed, loss_per_epoch, hipStream_t(s)); });
}
int agwn_mlp_grad(agwn_mlp* m, float const* pts, float const* target, int64_t n, float* grad, float* pred, float* loss)
{
    return guarded_mlp([&] { return agwn::mlp_grad(m, pts, target, n, grad, pred, loss); });
}
int agwn_mlp_predict_dev(agwn_mlp const* m, float const* pts, int64_t n, float* out, agwn_stream s)
{
    return guarded_mlp([&] { return agwn::mlp_predict_dev(m, pts, n, out, hipStream_t(s)); });
}
int agwn_mlp_set_surface(agwn_mlp* m, float const* verts, int64_t nv, int32_t const* idx, int64_t nidx, float eps,
                         float voxel)
{
    return guarded_mlp([&] { return agwn::mlp_set_surface(m, verts, nv, idx, nidx, eps, voxel); });
}
int agwn_training_samples_dev(float const* verts, int32_t const* idx, double const* cdf, int64_t ntri, float const* lo,
                              float const* hi, int64_t n, double near_frac, double band, double power, uint64_t seed,
                              float* out, agwn_stream s)
{
    return guarded_mlp([&] {
        return agwn::training_samples_dev(verts, idx, cdf, ntri, lo, hi, n, near_frac, band, power, seed, out,
                                          hipStream_t(s));
    });
}
int agwn_mlp_near_surface_dev(agwn_mlp const* m, float const* pts, int64_t n, int8_t* out, agwn_stream s)
{
    return guarded_mlp([&] { return agwn::mlp_near_surface_dev(m, pts, n, out, hipStream_t(s)); });
}
} // extern "C"
