// Learned absorbing boundary: canonical boundary samples, network inference,
// training (fused forward + backward with bf16 MFMA, FP32 accumulation,
// deterministic per-block partial gradients) and the Adam update.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "fdtd_internal.h"
#include "lbc_impl.h"

namespace fdtd {
namespace lbc {
namespace {

constexpr int MS = 72;     // transposed tiles: 64 rows of a round + 8
constexpr int DS = 40;     // dy rows (32 + 8)

// Samples of an nx x ny field window (row stride ld) -> out[npos][NS].
__global__ void __launch_bounds__(256) sample_kernel(int nx, int ny, int ld, const float *ez,
                                                     const float *hx, const float *hy, float *out)
{
    const int np = npos(nx, ny);
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    int e, t;
    pos_of(p, nx, ny, e, t);
    const int L = e < 2 ? ny : nx;
    const float *F[3] = {ez, hx, hy};
    auto acc = [&](int f, int i, int j) { return F[f][(long)i * ld + j]; };
    float *o = out + (long)p * NS;
    for (int oo = -1; oo <= 1; ++oo) {
        int tt = min(max(t + oo, 0), L - 1);
        for (int f = 0; f < 3; ++f)
            for (int d = 0; d < 3; ++d) o[sidx(f, d, oo)] = canon(acc, e, f, d, oo, tt, nx, ny);
    }
    o[S_EZ2] = canon(acc, e, 0, -1, 0, t, nx, ny);
    o[S_EZ1] = canon(acc, e, 0, -2, 0, t, nx, ny);
    o[S_EZ0] = canon(acc, e, 0, -3, 0, t, nx, ny);
}

// bf16 operand buffer from FP32 parameters (padding written as zero).
__global__ void __launch_bounds__(256) pack_kernel(const float *p, uint16_t *w)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= WBUF_ALL) return;
    if (q >= WBUF) {
        // the flag: does the hidden layers' output layer hold a non-zero
        // parameter (W3 and b3 are contiguous in the FP32 layout)?
        uint16_t any = 0;
        if (q == W_MLP)
            for (int k = P_W3; k < P_L; ++k) any |= p[k] != 0.0f;
        w[q] = any;
        return;
    }
    float v = 0.0f;
    if (q < W_W2) {
        int j = q / K1, k = q % K1;
        if (k < IN) v = p[P_W1 + j * IN + k];
    } else if (q < W_W3) {
        v = p[P_W2 + (q - W_W2)];
    } else if (q < W_W2T) {
        int j = (q - W_W3) / HID, k = (q - W_W3) % HID;
        if (j < OUT) v = p[P_W3 + j * HID + k];
    } else if (q < W_W3T) {
        int k = (q - W_W2T) / HID, j = (q - W_W2T) % HID;      // W2T[k][j] = W2[j][k]
        v = p[P_W2 + j * HID + k];
    } else if (q < W_FRM) {
        int k = (q - W_W3T) / 32, j = (q - W_W3T) % 32;        // W3T[k][j] = W3[j][k]
        if (j < OUT) v = p[P_W3 + j * HID + k];
    } else {
        // per-lane operands of the transposed network (inverse of frm_slot)
        const int e = q - W_FRM, lane = e / (FRM_OPS * 8), op = (e / 8) % FRM_OPS, vv = e % 8;
        const int c = lane & 15, g = lane >> 4;
        const int u = (2 * ((op - 16) % 2) + vv / 4) * 16 + 4 * g + vv % 4;
        if (op < 16) {
            const int i = 8 * (op % 4) + vv;
            if (i < NF) v = p[P_W1 + ((op / 4) * 16 + c) * IN + g * NF + i];
        } else if (op < 24) {
            v = p[P_W2 + (((op - 16) / 2) * 16 + c) * HID + u];
        } else if (c < OUT) {
            const int u3 = (2 * (op - 24) + vv / 4) * 16 + 4 * g + vv % 4;
            v = p[P_W3 + c * HID + u3];
        }
    }
    w[q] = f2bf(v);
}

// Raw features of row `row` (sample g of 4) from src: sample s of a row at
// offs[row] - s * tstride.
__device__ __forceinline__ void load_sample(const float *src, const long long *offs,
                                            long long tstride, int n, int row, int g,
                                            float (&v)[NF])
{
    if (row < n) {
        const float *q = src + offs[row] - (long long)g * tstride;
#pragma unroll
        for (int i = 0; i < NF; ++i) v[i] = q[i];
    } else {
#pragma unroll
        for (int i = 0; i < NF; ++i) v[i] = 0.0f;
    }
}

// Inference: y[row][3] (normalised outputs) and s[row] for n rows.
__global__ void __launch_bounds__(256) forward_kernel(int n, const float *src,
                                                      const long long *offs, long long tstride,
                                                      const float *p, const uint16_t *w, float *y,
                                                      float *s_out)
{
    __shared__ __attribute__((aligned(16))) uint16_t Xh[64 * XS];
    __shared__ __attribute__((aligned(16))) uint16_t A1[64 * AS];
    __shared__ __attribute__((aligned(16))) uint16_t A2[64 * AS];
    __shared__ float lin_s[32][OUT];
    const Net N{w, p};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, c = lane & 15, g = lane >> 4;
    const int r0 = wave * 16;
    // 32 samples per round: tile rows c and c + 8 hold sample 8 wave + (c & 7)
    // and its negation
    for (int base = blockIdx.x * 32; base < n; base += gridDim.x * 32) {
        const int row = base + wave * 8 + (c & 7);
        float v[NF];
        load_sample(src, offs, tstride, n, row, g, v);
        float s, inv;
        store_features(v, Xh, r0, s, inv);
        float lin[OUT];
        linear_out(p + P_L, v, inv, lin);
        if (g == 0 && c < 8) {
            if (row < n && s_out) s_out[row] = s;
#pragma unroll
            for (int d = 0; d < OUT; ++d) lin_s[wave * 8 + c][d] = lin[d];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        float z1[16], z2[16], yy[4];
        forward_tile<false>(N, Xh, A1, A2, r0, z1, z2, yy);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float yo = odd_out(yy[k]);
            const int rr = base + wave * 8 + g * 4 + k;
            if (g < 2 && c < OUT && rr < n) y[(long)rr * OUT + c] = lin_s[wave * 8 + g * 4 + k][c] + yo;
        }
        __syncthreads();
    }
}

struct TrainSmem {
    uint16_t Xh[64 * XS];
    uint16_t XhT[K1 * MS];
    uint16_t A1[64 * AS], A1T[HID * MS], A2[64 * AS], A2T[HID * MS];
    uint16_t DY[64 * DS], DYT[16 * MS];
    uint16_t DZ2[64 * AS], DZ2T[HID * MS], DZ1T[HID * MS];
    float th[32][4];
    float red[4][HID];
};

// One block: rounds of 32 samples as 64 MFMA rows (4 waves x 16; tile rows
// c and c + 8 hold a sample and its negation, for the odd network), forward
// + backward through bf16 MFMA; weight gradients accumulate in MFMA
// registers over the rounds (K = the 64 rows of a round); writes this
// block's partial gradient (NPARAM floats) and loss.
__global__ void __launch_bounds__(256) train_kernel(int n, const float *src,
                                                    const long long *offs, long long tstride,
                                                    const float *tgt, float norm, const float *p,
                                                    const uint16_t *w, float *part,
                                                    float *part_loss, float noise, uint32_t seed)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    TrainSmem &S = *reinterpret_cast<TrainSmem *>(smem_raw);
    const Net N{w, p};
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, c = lane & 15, g = lane >> 4;
    const int r0 = wave * 16;
    // zero the padding that no round writes: xh^T rows 108..127, dy columns 16..39
    for (int q = tid; q < (K1 - IN) * MS; q += 256) S.XhT[IN * MS + q] = 0;
    for (int q = tid; q < 64 * DS; q += 256) S.DY[q] = 0;
    __syncthreads();

    f32x4 gw1[8], gw2[4], gw3;
    for (int k = 0; k < 8; ++k) gw1[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int k = 0; k < 4; ++k) gw2[k] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    gw3 = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.0f, lsum = 0.0f;
    const float sc = 2.0f / norm;

    for (int base = blockIdx.x * 32; base < n; base += gridDim.x * 32) {
        // gather + scale (tile row c: sample wave*8 + (c & 7), age g)
        const int si = wave * 8 + (c & 7);
        const int row = base + si;
        float v[NF];
        load_sample(src, offs, tstride, n, row, g, v);
        float s, inv;
        store_features(v, S.Xh, r0, s, inv, S.XhT, MS, noise, (uint32_t)row, seed);
        // the hidden layers fit what the (fixed) linear path leaves
        float lin[OUT];
        linear_out(p + P_L, v, inv, lin);
        if (g == 0 && c < 8) {
            float wgt = (row < n && s > S_MIN) ? 1.0f : 0.0f;
#pragma unroll
            for (int d = 0; d < OUT; ++d) {
                float T = 0.0f, b = 0.0f;
                if (row < n) {
                    const float *nxt = src + offs[row] + tstride;
                    T = tgt ? tgt[(long)row * OUT + d]
                            : nxt[d == 2 ? S_EZ2 : (d == 1 ? S_EZ1 : S_EZ0)];
                    b = base_of(d, src + offs[row] - tstride);
                }
                S.th[si][d] = (T - b) * inv - lin[d];
            }
            S.th[si][3] = wgt;
        }
        __syncthreads();

        float z1[16], z2[16], y[4];
        forward_tile<true>(N, S.Xh, S.A1, S.A2, r0, z1, z2, y, S.A1T, S.A2T, MS);

        // loss gradient: e of sample wave*8 + (4g + k) & 7 through the odd
        // output; row dy = +-e * (2 / norm) / 2 (the f(x) and f(-x) branches)
        float dy[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float yo = odd_out(y[k]);
            const int rr = r0 + g * 4 + k;
            const int sj = wave * 8 + ((g * 4 + k) & 7);
            float e = 0.0f;
            if (c < OUT) e = (yo - S.th[sj][c]) * S.th[sj][3];
            if (g < 2) lsum += e * e;
            dy[k] = (g < 2 ? 0.5f : -0.5f) * (e * sc);
            gb3 += dy[k];
            uint16_t h = f2bf(dy[k]);
            S.DY[rr * DS + c] = h;
            S.DYT[c * MS + rr] = h;
        }
        __syncthreads();
        // da2 = dy W3 -> dz2
        {
            const bf16x8 a = ld8(S.DY + (r0 + c) * DS + g * 8);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                acc = mfma(a, ld8(w + W_W3T + (nt * 16 + c) * 32 + g * 8), acc);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    float dz = acc[k] * gelu_d(z2[nt * 4 + k]);
                    gb2[nt] += dz;
                    uint16_t h = f2bf(dz);
                    S.DZ2[(r0 + g * 4 + k) * AS + nt * 16 + c] = h;
                    S.DZ2T[(nt * 16 + c) * MS + r0 + g * 4 + k] = h;
                }
            }
        }
        __syncthreads();
        // da1 = dz2 W2 -> dz1
        {
            bf16x8 a[2];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) a[ks] = ld8(S.DZ2 + (r0 + c) * AS + ks * 32 + g * 8);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
                    acc = mfma(a[ks], ld8(w + W_W2T + (nt * 16 + c) * HID + ks * 32 + g * 8), acc);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    float dz = acc[k] * gelu_d(z1[nt * 4 + k]);
                    gb1[nt] += dz;
                    S.DZ1T[(nt * 16 + c) * MS + r0 + g * 4 + k] = f2bf(dz);
                }
            }
        }
        __syncthreads();
        // weight gradients over the round's 64 samples: wave owns hidden-unit
        // tile `wave` of W1 and W2 and input tile `wave` of W3
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 a1 = ld8(S.DZ1T + (wave * 16 + c) * MS + ks * 32 + g * 8);
#pragma unroll
            for (int kt = 0; kt < 8; ++kt)
                gw1[kt] = mfma(a1, ld8(S.XhT + (kt * 16 + c) * MS + ks * 32 + g * 8), gw1[kt]);
            const bf16x8 a2 = ld8(S.DZ2T + (wave * 16 + c) * MS + ks * 32 + g * 8);
#pragma unroll
            for (int kt = 0; kt < 4; ++kt)
                gw2[kt] = mfma(a2, ld8(S.A1T + (kt * 16 + c) * MS + ks * 32 + g * 8), gw2[kt]);
            const bf16x8 ay = ld8(S.DYT + c * MS + ks * 32 + g * 8);
            gw3 = mfma(ay, ld8(S.A2T + (wave * 16 + c) * MS + ks * 32 + g * 8), gw3);
        }
        __syncthreads();
    }

    float *P = part + (long)blockIdx.x * NPARAM;
#pragma unroll
    for (int kt = 0; kt < 8; ++kt)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int j = wave * 16 + g * 4 + k, kk = kt * 16 + c;
            if (kk < IN) P[P_W1 + j * IN + kk] = gw1[kt][k];
        }
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int k = 0; k < 4; ++k) P[P_W2 + (wave * 16 + g * 4 + k) * HID + kt * 16 + c] = gw2[kt][k];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = g * 4 + k;
        if (j < OUT) P[P_W3 + j * HID + wave * 16 + c] = gw3[k];
    }
    // L is fitted by least squares (fit_kernel), not by gradient: zero
    for (int q = tid; q < OUT * IN; q += 256) P[P_L + q] = 0.0f;
    // bias gradients and loss: lanes -> waves -> block, fixed order
    auto lane_sum = [](float x) {
        x += __shfl_xor(x, 16);
        x += __shfl_xor(x, 32);
        return x;
    };
    for (int nt = 0; nt < 4; ++nt) {
        float x = lane_sum(gb1[nt]);
        if (g == 0) S.red[wave][nt * 16 + c] = x;
    }
    __syncthreads();
    if (tid < HID) P[P_B1 + tid] = ((S.red[0][tid] + S.red[1][tid]) + S.red[2][tid]) + S.red[3][tid];
    __syncthreads();
    for (int nt = 0; nt < 4; ++nt) {
        float x = lane_sum(gb2[nt]);
        if (g == 0) S.red[wave][nt * 16 + c] = x;
    }
    __syncthreads();
    if (tid < HID) P[P_B2 + tid] = ((S.red[0][tid] + S.red[1][tid]) + S.red[2][tid]) + S.red[3][tid];
    __syncthreads();
    {
        float x = lane_sum(gb3);
        float l = lsum;
        for (int o = 1; o < 64; o <<= 1) l += __shfl_xor(l, o);
        if (g == 0) S.red[wave][c] = x;
        if (lane == 0) S.red[wave][16] = l;
    }
    __syncthreads();
    if (tid < OUT)
        P[P_B3 + tid] = ((S.red[0][tid] + S.red[1][tid]) + S.red[2][tid]) + S.red[3][tid];
    if (tid == 0 && part_loss)
        part_loss[blockIdx.x] = ((S.red[0][16] + S.red[1][16]) + S.red[2][16]) + S.red[3][16];
}

// Sum of the block partials (fixed order), Adam, bf16 operand copies.
__global__ void __launch_bounds__(256) adam_kernel(int nblocks, const float *part, float *p,
                                                   float *m, float *v, float lr, float b1,
                                                   float b2, float eps, float c1, float c2,
                                                   uint16_t *w, float *grad_out, int q0, int q1)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NPARAM) return;
    float gsum = 0.0f;
    for (int b = 0; b < nblocks; ++b) gsum += part[(long)b * NPARAM + q];
    if (grad_out) grad_out[q] = gsum;
    if (lr == 0.0f || q < q0 || q >= q1) return;
    const float mm = b1 * m[q] + (1.0f - b1) * gsum;
    const float vv = b2 * v[q] + (1.0f - b2) * gsum * gsum;
    m[q] = mm;
    v[q] = vv;
    const float val = p[q] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[q] = val;
    // the hidden layers' output turned non-zero: the time-stepping kernel
    // must evaluate them (the flag is only ever raised here; lbc_pack
    // recomputes it)
    if (q >= P_W3 && q < P_L && val != 0.0f) w[W_MLP] = 1;
    const uint16_t h = f2bf(val);
    const int fs = frm_slot(q);
    if (fs >= 0) w[fs] = h;
    if (q < P_B1) {
        const int j = q / IN, k = q % IN;
        w[W_W1 + j * K1 + k] = h;
    } else if (q >= P_W2 && q < P_B2) {
        const int j = (q - P_W2) / HID, k = (q - P_W2) % HID;
        w[W_W2 + j * HID + k] = h;
        w[W_W2T + k * HID + j] = h;
    } else if (q >= P_W3 && q < P_B3) {
        const int j = (q - P_W3) / HID, k = (q - P_W3) % HID;
        w[W_W3 + j * HID + k] = h;
        w[W_W3T + k * 32 + j] = h;
    }
}

// ------------------------------------------------------------ linear path
// Ridge least squares for the linear path L: with xh the normalised inputs
// (bf16-rounded when round_bf16, as the bf16 operands see them) and
// t = (target - base) / s the normalised residual target of a row,
//   G = sum_rows w xh xh^T (IN x IN),  R = sum_rows w xh t^T (IN x OUT),
//   L^T = (G + ridge I)^-1 R,  ridge = rho * sum_rows w,
// w = 1 for rows with s > S_MIN.  Sums in FP64.  Each block accumulates
// the IN x (IN + OUT) products plus the row count over its rows; the fit
// kernel reduces the partials in a fixed order and solves by Cholesky.
constexpr int GC = IN + OUT;            // columns of [G | R]
constexpr int GSZ = IN * GC + 1;        // + row count
constexpr int GPER = (IN * GC + 255) / 256;

__global__ void __launch_bounds__(256) gram_kernel(int n, const float *src, const long long *offs,
                                                   long long tstride, const float *tgt,
                                                   int round_bf16, float wpow, double *part)
{
    __shared__ float X[32][GC + 1];
    __shared__ float RW[32];            // sqrt of the row weight s^wpow
    const int tid = threadIdx.x;
    double acc[GPER];
    int ia[GPER], ib[GPER];
#pragma unroll
    for (int j = 0; j < GPER; ++j) {
        acc[j] = 0.0;
        const int e = tid + 256 * j;
        ia[j] = e < IN * GC ? e / GC : 0;
        ib[j] = e < IN * GC ? e % GC : 0;
    }
    double cnt = 0.0;
    // loader: thread t < 128 holds age t & 3 of row t >> 2 of the round
    const int lr = tid >> 2, la = tid & 3;
    for (int base = blockIdx.x * 32; base < n; base += gridDim.x * 32) {
        const int row = base + lr;
        if (tid < 128) {
            float v[NF];
            const bool ok = row < n;
            if (ok) {
                const float *q = src + offs[row] - (long long)la * tstride;
#pragma unroll
                for (int i = 0; i < NF; ++i) v[i] = q[i];
            } else {
#pragma unroll
                for (int i = 0; i < NF; ++i) v[i] = 0.0f;
            }
            float mx = 0.0f;
#pragma unroll
            for (int i = 0; i < NF; ++i) mx = fmaxf(mx, fabsf(v[i]));
            mx = fmaxf(mx, __shfl_xor(mx, 1));
            mx = fmaxf(mx, __shfl_xor(mx, 2));
            const bool use = ok && mx > S_MIN;
            const float inv = use ? 1.0f / mx : 0.0f;
#pragma unroll
            for (int i = 0; i < NF; ++i) {
                float x = v[i] * inv;
                if (round_bf16) x = bf2f(f2bf(x));
                X[lr][la * NF + i] = x;
            }
            if (la == 0) {
                RW[lr] = use ? (wpow == 0.0f ? 1.0f : powf(mx, 0.5f * wpow)) : 0.0f;
#pragma unroll
                for (int d = 0; d < OUT; ++d) {
                    float t = 0.0f;
                    if (use) {
                        const float *nxt = src + offs[row] + tstride;
                        const float T = tgt ? tgt[(long)row * OUT + d]
                                            : nxt[d == 2 ? S_EZ2 : (d == 1 ? S_EZ1 : S_EZ0)];
                        t = (T - base_of(d, src + offs[row] - tstride)) * inv;
                    }
                    X[lr][IN + d] = t;
                }
                if (use) cnt += (double)RW[lr] * (double)RW[lr];
            }
        }
        __syncthreads();
        for (int r = 0; r < 32; ++r) {
            const double wr = (double)RW[r] * (double)RW[r];
#pragma unroll
            for (int j = 0; j < GPER; ++j)
                acc[j] += wr * ((double)X[r][ia[j]] * (double)X[r][ib[j]]);
        }
        __syncthreads();
    }
    double *P = part + (long)blockIdx.x * GSZ;
#pragma unroll
    for (int j = 0; j < GPER; ++j) {
        const int e = tid + 256 * j;
        if (e < IN * GC) P[e] = acc[j];
    }
    // row count: loader threads (la == 0) -> block, fixed order
    __shared__ double cs[32];
    if (tid < 128 && la == 0) cs[lr] = cnt;
    __syncthreads();
    if (tid == 0) {
        double c = 0.0;
        for (int r = 0; r < 32; ++r) c += cs[r];
        P[IN * GC] = c;
    }
}

// Partials -> [G | R | count], fixed order over blocks.
__global__ void __launch_bounds__(256) gram_reduce_kernel(int nblocks, const double *part,
                                                          double *out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= GSZ) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[(long)b * GSZ + e];
    out[e] = s;
}

// One workgroup: Cholesky of G + ridge I in LDS (FP64, packed lower
// triangle), two triangular solves for the OUT right-hand sides, L written
// into the parameters.  status[0] = 0 on success, 1 + k when pivot k is not
// above 1e-13 of the largest diagonal entry (numerically singular: the
// system is then left unsolved and L is not written).
__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

__global__ void __launch_bounds__(256) fit_kernel(const double *gram, double rho, float *params,
                                                  int *status)
{
    __shared__ double A[IN * (IN + 1) / 2];
    __shared__ double B[IN][OUT];
    __shared__ int bad;
    const int tid = threadIdx.x;
    const double ridge = rho * gram[IN * GC];
    for (int i = 0; i < IN; ++i)
        for (int j = tid; j <= i; j += 256) A[tri(i, j)] = gram[i * GC + j] + (i == j ? ridge : 0.0);
    for (int e = tid; e < IN * OUT; e += 256) B[e / OUT][e % OUT] = gram[(e / OUT) * GC + IN + e % OUT];
    __shared__ double floor_s;
    if (tid == 0) {
        bad = 0;
        double m = 0.0;
        for (int i = 0; i < IN; ++i) m = fmax(m, gram[i * GC + i] + ridge);
        floor_s = 1e-13 * m;
    }
    __syncthreads();
    // right-looking Cholesky: A = C C^T, C lower
    for (int k = 0; k < IN; ++k) {
        if (tid == 0) {
            const double d = A[tri(k, k)];
            const bool ok = d > floor_s && d > 0.0;
            if (!ok && !bad) bad = 1 + k;
            A[tri(k, k)] = ok ? sqrt(d) : 1.0;
        }
        __syncthreads();
        const double dk = A[tri(k, k)];
        for (int i = k + 1 + tid; i < IN; i += 256) A[tri(i, k)] /= dk;
        __syncthreads();
        const int m = IN - k - 1;
        for (int e = tid; e < m * m; e += 256) {
            const int i = k + 1 + e / m, j = k + 1 + e % m;
            if (j <= i) A[tri(i, j)] -= A[tri(i, k)] * A[tri(j, k)];
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) status[0] = bad;
        return;
    }
    // C y = B, then C^T x = y (one thread per right-hand side)
    if (tid < OUT) {
        const int o = tid;
        for (int i = 0; i < IN; ++i) {
            double s = B[i][o];
            for (int j = 0; j < i; ++j) s -= A[tri(i, j)] * B[j][o];
            B[i][o] = s / A[tri(i, i)];
        }
        for (int i = IN - 1; i >= 0; --i) {
            double s = B[i][o];
            for (int j = i + 1; j < IN; ++j) s -= A[tri(j, i)] * B[j][o];
            B[i][o] = s / A[tri(i, i)];
        }
    }
    __syncthreads();
    for (int e = tid; e < OUT * IN; e += 256) params[P_L + e] = (float)B[e % IN][e / IN];
    if (tid == 0) status[0] = 0;
}

}  // namespace
}  // namespace lbc

int preload_lbc() { return preload_kernel(reinterpret_cast<const void *>(&lbc::pack_kernel)); }
}  // namespace fdtd

using namespace fdtd;
using namespace fdtd::lbc;

static hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

int lbc_param_count(void) { return NPARAM; }
int lbc_wbuf_elems(void) { return WBUF_ALL; }
size_t lbc_train_smem(void) { return sizeof(TrainSmem); }

int lbc_sample(int nx, int ny, int ld, const float *ez, const float *hx, const float *hy,
               float *out, void *stream)
{
    if (nx < 8 || ny < 8 || ld < ny) return fail("lbc_sample: window %dx%d ld=%d", nx, ny, ld);
    if (!ez || !hx || !hy || !out) return fail("lbc_sample: null pointer");
    const int np = npos(nx, ny);
    hipLaunchKernelGGL(sample_kernel, dim3((np + 255) / 256), dim3(256), 0, S(stream), nx, ny, ld,
                       ez, hx, hy, out);
    return hip_check(hipGetLastError(), "lbc_sample");
}

int lbc_pack(const float *params, uint16_t *wbuf, void *stream)
{
    if (!params || !wbuf) return fail("lbc_pack: null pointer");
    hipLaunchKernelGGL(pack_kernel, dim3((WBUF_ALL + 255) / 256), dim3(256), 0, S(stream), params,
                       wbuf);
    return hip_check(hipGetLastError(), "lbc_pack");
}

int lbc_forward(int n, const float *src, const long long *offs, long long tstride,
                const float *params, const uint16_t *wbuf, float *y, float *s, void *stream)
{
    if (n < 0) return fail("lbc_forward: n=%d", n);
    if (n == 0) return 0;
    if (!src || !offs || !params || !wbuf || !y) return fail("lbc_forward: null pointer");
    int blocks = (n + 31) / 32;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(forward_kernel, dim3(blocks), dim3(256), 0, S(stream), n, src, offs,
                       tstride, params, wbuf, y, s);
    return hip_check(hipGetLastError(), "lbc_forward");
}

int lbc_train_grad(int n, const float *src, const long long *offs, long long tstride,
                   const float *tgt, float norm, const float *params, const uint16_t *wbuf,
                   int nblocks, float *part, float *part_loss, float noise, unsigned seed,
                   void *stream)
{
    if (!(noise >= 0.0f)) return fail("lbc_train_grad: noise=%g", noise);
    if (n <= 0 || nblocks <= 0) return fail("lbc_train_grad: n=%d nblocks=%d", n, nblocks);
    if (!(norm > 0.0f)) return fail("lbc_train_grad: norm must be positive");
    if (!src || !offs || !params || !wbuf || !part) return fail("lbc_train_grad: null pointer");
    static thread_local bool attr = false;
    if (!attr) {
        int rc = hip_check(hipFuncSetAttribute((const void *)train_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)sizeof(TrainSmem)),
                           "lbc_train_grad: shared memory");
        if (rc) return rc;
        attr = true;
    }
    hipLaunchKernelGGL(train_kernel, dim3(nblocks), dim3(256), sizeof(TrainSmem), S(stream), n,
                       src, offs, tstride, tgt, norm, params, wbuf, part, part_loss, noise,
                       (uint32_t)seed);
    return hip_check(hipGetLastError(), "lbc_train_grad");
}

int lbc_adam(int nblocks, const float *part, float *params, float *m, float *v, int step,
             float lr, float beta1, float beta2, float eps, uint16_t *wbuf, float *grad_out,
             int first, int count, void *stream)
{
    if (first < 0 || count < 0) return fail("lbc_adam: first=%d count=%d", first, count);
    const int q0 = first, q1 = count ? first + count : NPARAM;
    if (nblocks <= 0 || !part) return fail("lbc_adam: nblocks=%d", nblocks);
    if (lr != 0.0f && (!params || !m || !v || !wbuf || step < 1))
        return fail("lbc_adam: params, m, v, wbuf and step >= 1 are required for an update");
    double c1 = 1.0 - __builtin_pow((double)beta1, step > 0 ? step : 1);
    double c2 = 1.0 - __builtin_pow((double)beta2, step > 0 ? step : 1);
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, S(stream), nblocks,
                       part, params, m, v, lr, beta1, beta2, eps, (float)c1, (float)c2, wbuf,
                       grad_out, q0, q1);
    return hip_check(hipGetLastError(), "lbc_adam");
}

int lbc_gram_elems(void) { return GSZ; }

int lbc_gram(int n, const float *src, const long long *offs, long long tstride, const float *tgt,
             int round_bf16, float wpow, int nblocks, double *part, void *stream)
{
    if (n <= 0 || nblocks <= 0) return fail("lbc_gram: n=%d nblocks=%d", n, nblocks);
    if (!src || !offs || !part) return fail("lbc_gram: null pointer");
    if (!(wpow >= 0.0f && wpow <= 4.0f)) return fail("lbc_gram: wpow=%g (0..4)", wpow);
    hipLaunchKernelGGL(gram_kernel, dim3(nblocks), dim3(256), 0, S(stream), n, src, offs, tstride,
                       tgt, round_bf16, wpow, part);
    return hip_check(hipGetLastError(), "lbc_gram");
}

int lbc_fit_linear(int nblocks, const double *part, double *gram, double rho, float *params,
                   int *status, void *stream)
{
    if (nblocks <= 0 || !part || !gram) return fail("lbc_fit_linear: nblocks=%d", nblocks);
    if (!(rho >= 0.0)) return fail("lbc_fit_linear: rho=%g", rho);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3((GSZ + 255) / 256), dim3(256), 0, S(stream),
                       nblocks, part, gram);
    int rc = hip_check(hipGetLastError(), "lbc_fit_linear: reduce");
    if (rc || !params) return rc;
    if (!status) return fail("lbc_fit_linear: status is required with params");
    hipLaunchKernelGGL(fit_kernel, dim3(1), dim3(256), 0, S(stream), gram, rho, params, status);
    return hip_check(hipGetLastError(), "lbc_fit_linear: solve");
}

}  // extern "C"
