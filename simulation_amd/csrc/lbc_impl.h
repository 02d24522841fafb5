// Learned absorbing boundary: network constants, bf16 helpers, the canonical
// boundary sample and the MFMA forward tile shared by the training and the
// time-stepping kernels.
//
// The network maps the last four canonical samples of an edge position
// (4 x 27 = 108 features: ez, normal-coupled H, tangential-coupled H at grid
// depths 2..4 x 3 tangential offsets) to the next-step ez of the outer 3
// cells:
//   s = max|x|, xh = bf16(x * (1/s))  (rows with s <= 1e-30 predict base),
//   f(xh) = W3 bf16(gelu(W2 bf16(gelu(W1 xh + b1)) + b2)) + b3,
//   gelu(z) = z sigma(1.702 z),
//   y = L xf + (f(xh) - f(-xh)) / 2   (odd: the exact boundary map is linear),
//   ez_next[d] = base[d] + s * y[d],
// with xf = x * (1/s) in FP32: the linear path L carries most of the
// correction and runs in FP32 on the vector ALUs (its least-squares
// coefficients cancel each other, and rounded to bf16 they turn the closed
// loop unstable); the hidden layers only correct what L leaves.
// base = the two-step absorbing boundary of fd1d/pycuda/fd1d_1_2.py:53-61
// applied to the three outer layers (depth d takes depth d+1's ez of two
// steps earlier): the network learns a correction to it, and a network with
// zero output is that classical boundary.  Inputs are the samples of grid
// depths 3..5 (see canon).  In a 16-row MFMA tile, rows r and r + 8 hold a
// sample and its negation.
// Matrix products run on v_mfma_f32_16x16x32_bf16 (bf16 operands, FP32
// accumulation).  No reference counterpart (the reference's boundaries are
// the two-step 1-D ABC and the Sullivan PML).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdtd {
namespace lbc {

constexpr int IN = 108, HID = 64, OUT = 3, NF = 27, HIST = 4;
// stored sample: the NF features, then ez of grid depths 2, 1, 0 at offset 0
// (training targets and baseline; never network inputs)
constexpr int NS = NF + 3, S_EZ2 = NF, S_EZ1 = NF + 1, S_EZ0 = NF + 2;
// rows whose scale s = max|x| is at most S_MIN count as all-zero (no
// prediction, no training weight): 1/s must stay finite
constexpr float S_MIN = 1e-30f;
constexpr int NPARAM = HID * IN + HID + HID * HID + HID + OUT * HID + OUT + OUT * IN;  // 11655
// flat FP32 parameter vector: W1 (64x108) b1 W2 (64x64) b2 W3 (3x64) b3, then
// the linear path L (3x108) from the inputs straight to the outputs
constexpr int P_W1 = 0, P_B1 = P_W1 + HID * IN, P_W2 = P_B1 + HID, P_B2 = P_W2 + HID * HID,
              P_W3 = P_B2 + HID, P_B3 = P_W3 + OUT * HID, P_L = P_B3 + OUT;
// bf16 operand buffer (uint16 elements): W1 [64][128] (k >= 108 zero),
// W2 [64][64], W3 [16][64] (rows >= 3 zero), W2^T [64][64], W3^T [64][32]
// (k >= 3 zero), then the per-lane operands of the transposed network of the
// time-stepping kernel (frm_slot); L is read in FP32 from the parameters
constexpr int K1 = 128;
constexpr int FRM_OPS = 26;     // per lane: W1 16 (tile mt, K step ks), W2 8, W3 2
constexpr int W_W1 = 0, W_W2 = W_W1 + HID * K1, W_W3 = W_W2 + HID * HID,
              W_W2T = W_W3 + 16 * HID, W_W3T = W_W2T + HID * HID, W_FRM = W_W3T + HID * 32,
              WBUF = W_FRM + 64 * FRM_OPS * 8;
// After the operands: one flag, 1 if the hidden layers' output (W3, b3) has
// a non-zero parameter, 0 if the network reduces to its linear path (the
// time-stepping kernel then skips the hidden layers); padded to 8 elements.
constexpr int W_MLP = WBUF, WBUF_ALL = WBUF + 8;

// Transposed network (positions as MFMA columns, hidden units as rows): lane
// (c = lane & 15, g = lane >> 4) holds, as A operands,
//   op mt*4 + ks      : W1[16 mt + c][27 g + 8 ks + v]        (8 ks + v < 27)
//   op 16 + mt*2 + ks : W2[16 mt + c][u(ks, g, v)]
//   op 24 + ks        : W3[c][u(ks, g, v)]                    (c < 3)
// v = 0..7, u(ks, g, v) = (2 ks + v / 4) 16 + 4 g + v % 4: the hidden unit a
// layer's accumulator leaves in lane (c, g) at K index 32 ks + 8 g + v of
// the next layer.  frm_slot maps a flat parameter index to its element of
// W_FRM (or -1: a bias, L, or W1's padding).
__host__ __device__ inline int frm_slot(int q)
{
    int lane, op, v;
    if (q < P_B1) {
        const int j = q / IN, k = q % IN, g = k / NF, i = k % NF;
        lane = (j % 16) + 16 * g; op = (j / 16) * 4 + i / 8; v = i % 8;
    } else if (q >= P_W2 && q < P_B2) {
        const int m = (q - P_W2) / HID, u = (q - P_W2) % HID, b = u / 16, r = u % 16;
        lane = (m % 16) + 16 * (r / 4); op = 16 + (m / 16) * 2 + b / 2; v = (b % 2) * 4 + r % 4;
    } else if (q >= P_W3 && q < P_B3) {
        const int o = (q - P_W3) / HID, u = (q - P_W3) % HID, b = u / 16, r = u % 16;
        lane = o + 16 * (r / 4); op = 24 + b / 2; v = (b % 2) * 4 + r % 4;
    } else {
        return -1;
    }
    return W_FRM + (lane * FRM_OPS + op) * 8 + v;
}

// LDS row strides (bf16 elements; multiples of 8 keep 16-byte alignment and
// spread 16 rows over distinct bank groups)
constexpr int XS = 136;   // xh rows (128 + 8)
constexpr int AS = 72;    // activation rows (64 + 8)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);          // nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t *p)
{
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    return __builtin_bit_cast(bf16x8, v);
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// GELU in the sigmoid form of Hendrycks & Gimpel (2016), z * sigma(1.702 z),
// with the hardware exp2 and reciprocal: a handful of instructions, where
// the erf form costs tens, and the result is rounded to bf16 anyway.
constexpr float GELU_A = 1.702f;
__device__ __forceinline__ float sigm(float u)
{
    return __builtin_amdgcn_rcpf(1.0f + __expf(-u));
}
__device__ __forceinline__ float gelu(float z) { return z * sigm(GELU_A * z); }
__device__ __forceinline__ float gelu_d(float z)
{
    const float sg = sigm(GELU_A * z);
    return sg * (1.0f + GELU_A * z * (1.0f - sg));
}

// Sample index of field f (0 ez, 1 normal-coupled H, 2 tangential-coupled
// H), depth d, tangential offset o in {-1, 0, 1}.
__host__ __device__ constexpr int sidx(int f, int d, int o) { return (f * 3 + d) * 3 + (o + 1); }

// Boundary positions: top (edge 0) and bottom (1) tau in [0, ny) own the
// corners; left (2) and right (3) tau in [3, nx-3).
__host__ __device__ inline int npos(int nx, int ny) { return 2 * ny + 2 * (nx - 6); }
__host__ __device__ inline void pos_of(int p, int nx, int ny, int &e, int &t)
{
    if (p < ny) { e = 0; t = p; }
    else if (p < 2 * ny) { e = 1; t = p - ny; }
    else if (p < 2 * ny + nx - 6) { e = 2; t = p - 2 * ny + 3; }
    else { e = 3; t = p - 2 * ny - (nx - 6) + 3; }
}
// grid cell of canonical depth d at tangential index t on edge e
__host__ __device__ inline void cell_of(int e, int d, int t, int nx, int ny, int &i, int &j)
{
    switch (e) {
    case 0: i = d; j = t; break;
    case 1: i = nx - 1 - d; j = t; break;
    case 2: i = t; j = d; break;
    default: i = t; j = ny - 1 - d; break;
    }
}

// Canonical value (field f, feature depth d, tangential offset o, tangential
// index t = clamped tau + o) of edge e, read through acc(field, i, j) with
// field 0 ez, 1 hx, 2 hy.  Feature depth d is grid depth d + 3: the three
// cells just inside the predicted ones, so the network never reads its own
// output back (a closed loop through its own predictions drifts).
// The four edges are mapped onto the top edge by the symmetries of the Yee
// update (x-mirror: hy -> -hy shifted one row; transpose: hx -> -hy^T,
// hy -> -hx^T); with D = d + 3:
//   top:    ez[D,t],      hn =  hy[D,t],       ht =  hx[D,t]
//   bottom: ez[nx-1-D,t], hn = -hy[nx-2-D,t],  ht =  hx[nx-1-D,t]
//   left:   ez[t,D],      hn = -hx[t,D],       ht = -hy[t,D]
//   right:  ez[t,ny-1-D], hn =  hx[t,ny-2-D],  ht = -hy[t,ny-1-D]
// ht is 0 at tangential offset +1 (ht is staggered half a cell along the
// edge: offsets -1 and 0 are the two H points beside the cell), so a
// prediction reaches one cell to either side.  H in the last row or column
// reads as 0: the H update never runs there (it stops at nx-2, ny-2), so a
// window cut from a larger grid (training data) must not see it either.
constexpr int FD0 = 3;      // grid depth of feature depth 0
template <class Acc>
__device__ __forceinline__ float canon(const Acc &acc, int e, int f, int d, int o, int t, int nx,
                                       int ny)
{
    if (f == 2 && o == 1) return 0.0f;
    const int D = d + FD0;
    int fld, i, j;
    float sg = 1.0f;
    switch (e) {
    case 0: i = D; j = t; fld = f == 0 ? 0 : (f == 1 ? 2 : 1); break;
    case 1:
        j = t;
        if (f == 0) { fld = 0; i = nx - 1 - D; }
        else if (f == 1) { fld = 2; i = nx - 2 - D; sg = -1.0f; }
        else { fld = 1; i = nx - 1 - D; }
        break;
    case 2: i = t; j = D; fld = f == 0 ? 0 : (f == 1 ? 1 : 2); sg = f == 0 ? 1.0f : -1.0f; break;
    default:
        i = t;
        if (f == 0) { fld = 0; j = ny - 1 - D; }
        else if (f == 1) { fld = 1; j = ny - 2 - D; }
        else { fld = 2; j = ny - 1 - D; sg = -1.0f; }
        break;
    }
    if (fld != 0 && (i == nx - 1 || j == ny - 1)) return 0.0f;
    const float v = acc(fld, i, j);
    return sg < 0.0f ? -v : v;
}

// Baseline of output d (grid depth d) from the sample of age 1 (the state
// before the current one): the two-step chain ez_d(t+1) = ez_{d+1}(t-1).
__device__ __forceinline__ float base_of(int d, const float *s1)
{
    return d == 2 ? s1[sidx(0, 0, 0)] : (d == 1 ? s1[S_EZ2] : s1[S_EZ1]);
}

// Odd output of sample 4g + k (g < 2) of a tile from the accumulator-layout
// y[k] of rows 4g + k and 4g + k + 8 (lane ^ 32): (f(x) - f(-x)) / 2.  All
// lanes must call it; lanes with g >= 2 get the same value.
__device__ __forceinline__ float odd_out(float yk)
{
    const int g = (threadIdx.x & 63) >> 4;
    const float other = __shfl_xor(yk, 32);
    return g < 2 ? 0.5f * (yk - other) : 0.5f * (other - yk);
}

struct Net {
    const uint16_t *w;      // bf16 operand buffer (WBUF_ALL elements)
    const float *p;         // FP32 parameters (biases and L read from here)
};

// Linear path of the sample in tile row r: lane (r, g) holds the raw
// features of age g in v (scale 1/s in inv); every lane of the row gets
// lin[o] = sum_k L[o][k] x[k] inv (FP32: per-lane sums over the lane's 27
// features in order, then the four ages as (g0 + g1) + (g2 + g3)).
__device__ __forceinline__ void linear_out(const float *Lp, const float (&v)[NF], float inv,
                                           float (&lin)[OUT])
{
    const int g = (threadIdx.x & 63) >> 4;
#pragma unroll
    for (int o = 0; o < OUT; ++o) {
        const float *w = Lp + o * IN + g * NF;
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < NF; ++i) a = __builtin_fmaf(w[i], v[i] * inv, a);
        a = a + __shfl_xor(a, 16);
        a = a + __shfl_xor(a, 32);
        lin[o] = a;
    }
}

// Scales and stores one 16-row tile's features.  Lane (r = lane&15,
// g = lane>>4) holds the 27 features of sample g of row r in v[]; writes xh
// to Xh[(r0 + r) * XS + 27 g + i] (and columns 108..127 as zero).  Returns
// the row scale s and 1/s (0 when s <= S_MIN) to every lane of the row.
// Counter-based uniform noise in [-1, 1) (integer hash of the arguments).
__device__ __forceinline__ float hash_unit(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ c * 0xC2B2AE3Du;
    h ^= h >> 15; h *= 0x2C1B3C6Du; h ^= h >> 12; h *= 0x297A2D39u; h ^= h >> 15;
    return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;
}

// Rows r and r + 8 of a tile hold a sample and its negation (the caller
// loads sample r & 7 into both).  With noise > 0 (training only) each
// normalised input gets + noise * u, u uniform in [-1, 1) from
// hash_unit(key, feature index, seed), negated with the sample.
__device__ __forceinline__ void store_features(const float (&v)[NF], uint16_t *Xh, int r0,
                                               float &s, float &inv, uint16_t *XhT = nullptr,
                                               int MS = 0, float noise = 0.0f,
                                               uint32_t key = 0, uint32_t seed = 0)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const bool neg = r >= 8;
    float mx = 0.0f;
#pragma unroll
    for (int i = 0; i < NF; ++i) mx = fmaxf(mx, fabsf(v[i]));
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    s = mx;
    inv = mx > S_MIN ? 1.0f / mx : 0.0f;
    uint16_t *row = Xh + (r0 + r) * XS;
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        float x = v[i] * inv;
        if (noise > 0.0f && inv > 0.0f) x = x + noise * hash_unit(key, g * NF + i, seed);
        uint16_t b = f2bf(neg ? -x : x);
        row[g * NF + i] = b;
        if (XhT) XhT[(g * NF + i) * MS + r0 + r] = b;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) row[IN + g * 5 + q] = 0;
}

// Forward pass of the 16 rows r0..r0+15 held in Xh (one wave).  Outputs in
// the MFMA accumulator layout: lane (c = lane&15, g = lane>>4) holds
// z1/z2[nt*4+v] of row r0 + 4g + v, hidden unit 16 nt + c, and y[v] of row
// r0 + 4g + v, output c (valid for c < 3).  a1/a2 (bf16) are left in A1/A2
// rows; with A1T/A2T also stored transposed (column = row index).
template <bool KEEP>
__device__ __forceinline__ void forward_tile(const Net &N, const uint16_t *Xh, uint16_t *A1,
                                             uint16_t *A2, int r0, float (&z1)[16],
                                             float (&z2)[16], float (&y)[4],
                                             uint16_t *A1T = nullptr, uint16_t *A2T = nullptr,
                                             int MS = 0)
{
    const int lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
    const uint16_t *W1 = N.w + W_W1, *W2 = N.w + W_W2, *W3 = N.w + W_W3;
    // layer 1: [16 x 128] x [128 x 64]
    bf16x8 xa[4];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) xa[ks] = ld8(Xh + (r0 + c) * XS + ks * 32 + g * 8);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = mfma(xa[ks], ld8(W1 + (nt * 16 + c) * K1 + ks * 32 + g * 8), acc);
        const float b = N.p[P_B1 + nt * 16 + c];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float z = acc[v] + b;
            z1[nt * 4 + v] = z;
            uint16_t h = f2bf(gelu(z));
            A1[(r0 + g * 4 + v) * AS + nt * 16 + c] = h;
            if (KEEP) A1T[(nt * 16 + c) * MS + r0 + g * 4 + v] = h;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 2: [16 x 64] x [64 x 64]
    bf16x8 aa[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) aa[ks] = ld8(A1 + (r0 + c) * AS + ks * 32 + g * 8);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
            acc = mfma(aa[ks], ld8(W2 + (nt * 16 + c) * HID + ks * 32 + g * 8), acc);
        const float b = N.p[P_B2 + nt * 16 + c];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            float z = acc[v] + b;
            z2[nt * 4 + v] = z;
            uint16_t h = f2bf(gelu(z));
            A2[(r0 + g * 4 + v) * AS + nt * 16 + c] = h;
            if (KEEP) A2T[(nt * 16 + c) * MS + r0 + g * 4 + v] = h;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 3: [16 x 64] x [64 x 16] (outputs 3..15 are zero rows of W3)
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
        acc = mfma(ld8(A2 + (r0 + c) * AS + ks * 32 + g * 8), ld8(W3 + c * HID + ks * 32 + g * 8),
                   acc);
    const float b3 = c < OUT ? N.p[P_B3 + c] : 0.0f;
#pragma unroll
    for (int v = 0; v < 4; ++v) y[v] = acc[v] + b3;
}

}  // namespace lbc
}  // namespace fdtd
