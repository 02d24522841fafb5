// Learned-boundary frame kernel.
//
// A learned-boundary launch of K leapfrog steps is the temporally blocked TM
// launch (no PML, PEC edges) over the whole grid followed by this kernel,
// which recomputes the frame of depth K+3 around the grid with the network
// applied at every step.  Why K+3: the learned update differs from the PEC
// update only in ez of the outer three cells, and a difference moves at most
// one cell inward per step (E reaches depth K+1 after K steps; on the bottom
// and right edges H at depth d takes ez at depth d-1, so H reaches K+2):
// after K steps the blocked launch is exact at depth >= K+3.
//
// Each workgroup owns one 32-cell chunk of one edge's frame (depth K+3) and
// evolves, in LDS, the chunk's dependency region.  Per step the valid part
// of a region shrinks by one cell on a cut side, except along an edge in the
// direction of increasing index, where it shrinks by two: H at tau takes ez
// at tau+1, whose prediction reads the samples at tau+2 (Yee staggering).
// The samples handed to the next launch reach one cell beyond the chunk and
// down to grid depth 6 (the bottom edge's H between depths 5 and 6).
// Regions: depth 2K+6 (3K+6 for the two corner chunks of the top edge, whose
// side-edge positions run down into the cut), widened along the edge by K+1
// cells before and 2K+1 after the chunk.  Per step: network for every boundary position
// whose cells lie in the region (features = the position's last four
// canonical samples, kept in an LDS ring), D and E update, ez of the
// predicted cells replaced, H update, new samples pushed.
//
// The network runs transposed, D^T = W X^T: the 16 MFMA columns are 8
// positions and their negations (the odd network), the rows are hidden
// units.  Lane (c, g) of a wave supplies column c's features of age g, so
// the input operand is built from the lane's own ring reads, and a layer's
// accumulator (hidden units 16 mt + 4 g + v of column c) is, after GELU and
// rounding, the next layer's input operand once the next layer's weight
// columns are permuted to match (unit (2 ks + v/4) 16 + 4 g + v%4 at K index
// 32 ks + 8 g + v): no LDS between the layers.  The layer weights sit in
// registers for the whole kernel; the FP32 linear path and the biases in
// LDS.  b3 is left out: the odd combination cancels it.  The ring enters
// and leaves through a global history buffer (four samples per position,
// newest first; ping-ponged between launches).  Expressions follow
// fd2d/program/fd2d_3_3.py:68-98 (free space) with the lossy E update of
// fd2d/pycuda/fd2d_3_4.py:181-182, operand for operand, no contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fdtd_internal.h"
#include "lbc_impl.h"

namespace fdtd {
namespace lbc {
namespace {

constexpr int CH = 32;                        // chunk length along the edge (at 2048^2:
                                              // 256 workgroups, one per CU)
constexpr int KMAX = 10;                      // steps per launch
constexpr int NT = 512;                       // threads per workgroup (8 waves)
// region cells: corner chunks (3K+6) x (CH + 2K + 1), others at most
// (2K+6) x (CH + 3K + 2) (30 x 49 and 22 x 58 at K = 8, 36 x 53 and
// 26 x 64 at K = 10)
constexpr int RMAX_C = (3 * KMAX + 6) * (CH + 2 * KMAX + 1);
constexpr int RMAX_E = (2 * KMAX + 6) * (CH + 3 * KMAX + 2);
constexpr int RMAX = ((RMAX_C > RMAX_E ? RMAX_C : RMAX_E) + 3) / 4 * 4;   // keeps the ring 16-byte aligned
constexpr int PMAX = 136;                     // positions per workgroup (130 at K = 10)

struct FrameSmem {
    float dz[RMAX], iz[RMAX], ez[RMAX], hx[RMAX], hy[RMAX], na[RMAX], nb[RMAX];
    float ring[PMAX][HIST][NS];
    float pred[PMAX][OUT];
    float L[OUT][IN];       // linear path (FP32)
    float b1[HID], b2[HID];
    int pos[PMAX];          // edge << 24 | tau
    int own[PMAX];          // global history index, or -1 when not owned
    // where each sample value comes from: bits 0-11 region cell, 12-13 field
    // (0 ez, 1 hx, 2 hy, 3 constant 0), bit 15 negate
    uint16_t src[PMAX][NS];
};

static_assert(sizeof(FrameSmem) + 64 <= 160 * 1024, "frame kernel LDS above 160 KiB");
static_assert(CH + 3 * KMAX + 2 + 2 * (3 * KMAX + 3) <= PMAX && PMAX % 8 == 0,
              "positions of a frame workgroup above PMAX");

// A workgroup's positions and sample sources depend only on the grid shape
// and the depth: built once (frame_setup_kernel) into global memory.
struct alignas(16) FrameTab {
    int np, pad[3];
    int nact[KMAX + 4];     // nact[l]: positions needed at level l (a prefix)
    int pos[PMAX], own[PMAX];
    uint16_t src[PMAX][NS];
};
static_assert(sizeof(FrameTab) % 16 == 0, "table stride");

struct Geo {
    int ri0, rj0, RH, RW;           // region origin and size (global)
    int oi0, oi1, oj0, oj1;         // owned output cells
};

__device__ inline int gidx(int e, int t, int nx, int ny)
{
    switch (e) {
    case 0: return t;
    case 1: return ny + t;
    case 2: return 2 * ny + t - 3;
    default: return 2 * ny + (nx - 6) + t - 3;
    }
}

// Layer weights of one lane in MFMA A-operand form (W_FRM of the bf16
// buffer, in frm_slot's order).
struct LaneW {
    bf16x8 w1[4][4];        // hidden tile mt, K step ks
    bf16x8 w2[4][2];
    bf16x8 w3[2];
};

__device__ __forceinline__ void load_lane_w(LaneW &W, const uint16_t *wbuf, int lane)
{
    const uint16_t *b = wbuf + W_FRM + lane * FRM_OPS * 8;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) W.w1[mt][ks] = ld8(b + (mt * 4 + ks) * 8);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) W.w2[mt][ks] = ld8(b + (16 + mt * 2 + ks) * 8);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) W.w3[ks] = ld8(b + (24 + ks) * 8);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// GELU of the accumulators of 4 hidden tiles (+ bias), rounded to bf16 and
// packed as the next layer's two K steps.
__device__ __forceinline__ void act_pack(const f32x4 (&z)[4], const float *bias, int g,
                                         bf16x8 (&out)[2])
{
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        bf16x2 h[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int mt = 2 * ks + (u >> 1), v0 = 2 * (u & 1);
            f32x2 a;
            a.x = gelu(z[mt][v0] + bias[mt * 16 + 4 * g + v0]);
            a.y = gelu(z[mt][v0 + 1] + bias[mt * 16 + 4 * g + v0 + 1]);
            h[u] = __builtin_convertvector(a, bf16x2);
        }
        out[ks] = __builtin_bit_cast(bf16x8, h);
    }
}

// Region and owned cells of workgroup b (kind: 0 top, 1 bottom, 2 left,
// 3 right chunk; q: chunk index along its edge).
__device__ inline void block_geo(int b, int nx, int ny, int K, Geo &G, int &kind, int &q)
{
    const int W = K + 3;
    const int nct = (ny + CH - 1) / CH;
    const int ncl = (nx - 2 * W + CH - 1) / CH;
    if (b < nct) { kind = 0; q = b; }
    else if (b < 2 * nct) { kind = 1; q = b - nct; }
    else if (b < 2 * nct + ncl) { kind = 2; q = b - 2 * nct; }
    else { kind = 3; q = b - 2 * nct - ncl; }
    int R0 = 2 * K + 6;
    if (kind < 2) {
        const int c0 = q * CH, c1 = min(ny, c0 + CH);
        if (kind == 0 && (q == 0 || q == nct - 1)) R0 = 3 * K + 6;
        G.rj0 = max(0, c0 - K - 1);
        G.RW = min(ny, c1 + 2 * K + 1) - G.rj0;
        G.RH = R0;
        G.ri0 = kind == 0 ? 0 : nx - R0;
        G.oj0 = c0; G.oj1 = c1;
        G.oi0 = kind == 0 ? 0 : nx - W;
        G.oi1 = kind == 0 ? W : nx;
    } else {
        const int r0 = W + q * CH, r1 = min(nx - W, r0 + CH);
        G.ri0 = max(0, r0 - K - 1);
        G.RH = min(nx, r1 + 2 * K + 1) - G.ri0;
        G.RW = R0;
        G.rj0 = kind == 2 ? 0 : ny - R0;
        G.oi0 = r0; G.oi1 = r1;
        G.oj0 = kind == 2 ? 0 : ny - W;
        G.oj1 = kind == 2 ? W : ny;
    }
}

// One workgroup per frame workgroup: its position list (thread 0; at most
// CH + 3K + 2 + 2 (3K + 3) positions: 112 at K = 8, 130 at K = 10) and
// sample-source table.
__global__ void __launch_bounds__(NT) frame_setup_kernel(int nx, int ny, int K, FrameTab *tabs)
{
    __shared__ FrameTab S;
    FrameTab &T = tabs[blockIdx.x];
    const int tid = threadIdx.x;
    const int W = K + 3;
    const int nct = (ny + CH - 1) / CH;
    Geo G;
    int kind, q;
    block_geo(blockIdx.x, nx, ny, K, G, kind, q);
    const int RH = G.RH, RW = G.RW;
    // ---------------------------------------------------------- positions
    // (thread 0 builds the list; at most CH + 3K + 2 + 2 (3K + 3) <= PMAX)
    if (tid == 0) {
        int n = 0;
        auto add = [&](int e, int t0, int t1, int w0, int w1) {
            for (int t = t0; t < t1 && n < PMAX; ++t) {
                S.pos[n] = (e << 24) | t;
                S.own[n] = (t >= w0 && t < w1) ? gidx(e, t, nx, ny) : -1;
                ++n;
            }
        };
        const int ilo = G.ri0, ihi = G.ri0 + RH, jlo = G.rj0, jhi = G.rj0 + RW;
        if (kind < 2) {
            // this edge's positions in the region; owned: the chunk
            add(kind, jlo, jhi, G.oj0, G.oj1);
            // side-edge positions whose cells lie in the region; the first and
            // last chunks own those the side chunks do not (rows < W or >= nx-W)
            const int s0 = kind == 0 ? 3 : max(3, ilo), s1 = kind == 0 ? min(ihi, nx - 3) : nx - 3;
            const int w0 = kind == 0 ? 3 : nx - W, w1 = kind == 0 ? W : nx - 3;
            if (jlo <= 2) add(2, s0, s1, w0, q == 0 ? w1 : w0);
            if (jhi >= ny - 2) add(3, s0, s1, w0, q == nct - 1 ? w1 : w0);
        } else {
            add(kind, max(3, ilo), min(nx - 3, ihi), G.oi0, G.oi1);
            const int t0 = kind == 2 ? 0 : jlo, t1 = kind == 2 ? jhi : ny;
            if (ilo <= 2) add(0, t0, t1, 0, -1);
            if (ihi >= nx - 2) add(1, t0, t1, 0, -1);
        }
        S.np = n;
    }

    __syncthreads();
    const int np = S.np;
    // Levels at which a position is needed.  Along an edge the valid part of
    // a region shrinks per step by one cell at a cut on the low side and by
    // two at a cut on the high side, so position t (region range [lo, hi)
    // in its edge's coordinate; a side at the grid's edge is no cut) has a
    // valid prediction up to level min(t - lo, (hi - t) / 2); it is
    // evaluated one level beyond (margin).  Positions sorted by that level,
    // descending (counting sort, stable): level l works on a prefix.
    __shared__ int key[PMAX], cnt[KMAX + 2], at[KMAX + 2];
    __shared__ int tpos[PMAX], town[PMAX];
    for (int k = tid; k < np; k += NT) {
        const int e = S.pos[k] >> 24, t = S.pos[k] & 0xFFFFFF;
        const int lo = e < 2 ? G.rj0 : G.ri0, hi = e < 2 ? G.rj0 + RW : G.ri0 + RH;
        const int Lg = e < 2 ? ny : nx;
        int lm = K;
        if (lo > 0) lm = min(lm, t - lo + 1);
        if (hi < Lg) lm = min(lm, (hi - t) / 2 + 1);
        key[k] = max(1, lm);
        tpos[k] = S.pos[k];
        town[k] = S.own[k];
    }
    __syncthreads();
    if (tid == 0) {
        for (int l = 0; l <= K + 1; ++l) cnt[l] = 0;
        for (int k = 0; k < np; ++k) ++cnt[key[k]];
        int o = 0;
        for (int l = K; l >= 1; --l) { at[l] = o; o += cnt[l]; }
        for (int k = 0; k < np; ++k) {
            const int d = at[key[k]]++;
            S.pos[d] = tpos[k];
            S.own[d] = town[k];
        }
        int n = 0;
        for (int l = K; l >= 1; --l) {
            n += cnt[l];
            T.nact[l] = n;
        }
        T.nact[0] = np;
    }
    __syncthreads();
    // sample sources, resolved once: canon() run on an accessor that returns
    // the address instead of the value (clamped into the region: positions
    // near a cut side read stale neighbours, within the part of the region
    // that is no longer valid anyway)
    for (int k = tid; k < np * NS; k += NT) {
        const int p = k / NS, i = k % NS;
        const int e = S.pos[p] >> 24, t = S.pos[p] & 0xFFFFFF;
        int code = -1;                       // set by the accessor
        auto where = [&](int f, int ii, int jj) {
            const int r = min(max(ii - G.ri0, 0), RH - 1), cc = min(max(jj - G.rj0, 0), RW - 1);
            code = (f << 12) | (r * RW + cc);
            return 1.0f;
        };
        float sg;
        if (i < NF) {
            const int f = i / 9, d = (i / 3) % 3, o = i % 3 - 1;
            const int L = e < 2 ? ny : nx;
            const int tt = min(max(t + o, 0), L - 1);
            sg = canon(where, e, f, d, o, tt, nx, ny);
        } else {
            sg = canon(where, e, 0, NF - 1 - i, 0, t, nx, ny);   // S_EZ2, S_EZ1, S_EZ0
        }
        uint16_t v;
        if (code < 0 || sg == 0.0f) v = (uint16_t)(3 << 12);
        else v = (uint16_t)(code | (sg < 0.0f ? 0x8000 : 0));
        T.src[p][i] = v;
    }
    for (int k = tid; k < np; k += NT) {
        T.pos[k] = S.pos[k];
        T.own[k] = S.own[k];
    }
    if (tid == 0) T.np = np;
}

__global__ void __launch_bounds__(NT) frame_kernel(LbcFrameParams P)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    FrameSmem &S = *reinterpret_cast<FrameSmem *>(smem_raw);
    const int nx = P.nx, ny = P.ny, K = P.K, W = K + 3;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, c = lane & 15, g = lane >> 4;
    constexpr int NW = NT / 64;

    Geo G;
    int kind, q;
    block_geo(blockIdx.x, nx, ny, K, G, kind, q);
    const FrameTab &T = reinterpret_cast<const FrameTab *>(P.tab)[blockIdx.x];
    const int RH = G.RH, RW = G.RW, NC = RH * RW;

    // ---------------------------------------------------------- tables
    const int np = T.np;
    __shared__ int nact_s[KMAX + 4];
    if (tid <= K) nact_s[tid] = T.nact[tid];
    for (int k = tid; k < np; k += NT) {
        S.pos[k] = T.pos[k];
        S.own[k] = T.own[k];
    }
    {
        const uint32_t *s32 = reinterpret_cast<const uint32_t *>(&T.src[0][0]);
        uint32_t *d32 = reinterpret_cast<uint32_t *>(&S.src[0][0]);
        for (int k = tid; k < np * NS / 2; k += NT) d32[k] = s32[k];
    }
    // ---------------------------------------------------------- load
    for (int k = tid; k < RH * RW; k += NT) {
        const int r = k / RW, cc = k % RW;
        const long o = (long)(G.ri0 + r) * ny + (G.rj0 + cc);
        S.dz[k] = P.dz_i[o];
        S.iz[k] = P.iz_i ? P.iz_i[o] : 0.0f;
        S.hx[k] = P.hx_i[o];
        S.hy[k] = P.hy_i[o];
        S.na[k] = P.naz[o];
        S.nb[k] = P.nbz ? P.nbz[o] : 0.0f;
        S.ez[k] = 0.0f;
    }
    // history ring: the sample of level l (l = -3..K) in slot (-l) & 3, so
    // the history's ages 0..3 (levels 0, -1, -2, -3) land in slots 0..3: one
    // contiguous copy of 4 x NS floats per position (positions read from the
    // table in global memory: no barrier between the loads)
    for (int k = tid; k < np * (HIST * NS / 4); k += NT) {
        const int p = k / (HIST * NS / 4), j = k % (HIST * NS / 4);
        const int ps = T.pos[p], e = ps >> 24, t = ps & 0xFFFFFF;
        const float4 v = reinterpret_cast<const float4 *>(P.hist_in + (long)gidx(e, t, nx, ny) * HIST * NS)[j];
        reinterpret_cast<float4 *>(&S.ring[p][0][0])[j] = v;
    }

    for (int k = tid; k < OUT * IN; k += NT) S.L[k / IN][k % IN] = P.params[P_L + k];
    for (int k = tid; k < HID; k += NT) {
        S.b1[k] = P.params[P_B1 + k];
        S.b2[k] = P.params[P_B2 + k];
    }
    LaneW LW;
    load_lane_w(LW, P.w, lane);
    // a network whose hidden layers' output layer is zero adds exactly
    // nothing (y = 0, so the odd part 0.5 (y - y') is +0, as computed): skip
    // those layers
    const bool mlp = P.w[W_MLP] != 0;
    // cell loops without divisions: thread tid starts at cell tid and steps
    // by NT = DR rows + DC columns
    const int DR = NT / RW, DC = NT % RW;
    const int r_start = tid / RW, c_start = tid % RW;
    __syncthreads();

    for (int lev = 1; lev <= K; ++lev) {
        // ------------------------------------------------------ network
        // tiles of 8 positions, tile q on wave q % NW; no barrier before the
        // D/E update, which touches neither the ring nor pred
        const int na = nact_s[lev];
        for (int tb = wave * 8; tb < na; tb += NW * 8) {
            const int q = tb + (c & 7);
            float x[NF];
            if (q < na) {
                const float *src = S.ring[q][(g + 1 - lev) & 3];   // level lev - 1 - g
#pragma unroll
                for (int i = 0; i < NF; ++i) x[i] = src[i];
            } else {
#pragma unroll
                for (int i = 0; i < NF; ++i) x[i] = 0.0f;
            }
            float mx = 0.0f;
#pragma unroll
            for (int i = 0; i < NF; ++i) mx = fmaxf(mx, fabsf(x[i]));
            mx = fmaxf(mx, __shfl_xor(mx, 16));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            const float inv = mx > S_MIN ? 1.0f / mx : 0.0f;
            // FP32 linear path: lanes c and c + 8 hold the same position and
            // split its 27 features of age g (0..13 and 14..26)
            float lin[OUT];
            {
                const int h0 = c < 8 ? 0 : 14;
                float xs[14];
#pragma unroll
                for (int i = 0; i < 14; ++i) {
                    const float lo = x[i], hi = i + 14 < NF ? x[i + 14] : 0.0f;
                    xs[i] = (c < 8 ? lo : hi) * inv;
                }
#pragma unroll
                for (int o = 0; o < OUT; ++o) {
                    const float *w = &S.L[o][g * NF + h0];
                    float a = 0.0f;
#pragma unroll
                    for (int i = 0; i < 14; ++i)
                        if (i < 13 || c < 8) a = __builtin_fmaf(w[i], xs[i], a);
                    a = a + __shfl_xor(a, 8);
                    a = a + __shfl_xor(a, 16);
                    a = a + __shfl_xor(a, 32);
                    lin[o] = a;
                }
            }
            f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
            if (mlp) {
            // layer 1 input: column c < 8 the features, c >= 8 their negation
            const float sc = c < 8 ? inv : -inv;
            bf16x8 xb[4];
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                bf16x2 h[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i0 = 8 * ks + 2 * u;
                    f32x2 a;
                    a.x = i0 < NF ? x[i0] * sc : 0.0f;
                    a.y = i0 + 1 < NF ? x[i0 + 1] * sc : 0.0f;
                    h[u] = __builtin_convertvector(a, bf16x2);
                }
                xb[ks] = __builtin_bit_cast(bf16x8, h);
            }
            f32x4 z[4];
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                z[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < 4; ++ks) z[mt] = mfma(LW.w1[mt][ks], xb[ks], z[mt]);
            }
            bf16x8 a1[2];
            act_pack(z, S.b1, g, a1);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                z[mt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) z[mt] = mfma(LW.w2[mt][ks], a1[ks], z[mt]);
            }
            bf16x8 a2[2];
            act_pack(z, S.b2, g, a2);
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) y = mfma(LW.w3[ks], a2[ks], y);
            }   // mlp
            // outputs v = 0..2 of column c in lanes g = 0: odd part from the
            // column pair (c, c + 8)
#pragma unroll
            for (int v = 0; v < OUT; ++v) {
                const float other = __shfl_xor(y[v], 8);
                if (g == 0 && c < 8 && q < na) {
                    const float yo = 0.5f * (y[v] - other);
                    const float b = base_of(v, S.ring[q][(2 - lev) & 3]);
                    S.pred[q][v] = mx > S_MIN ? b + mx * (lin[v] + yo) : b;
                }
            }
        }
        // ------------------------------------------------------ D, E update
        for (int k = tid, r = r_start, cc = c_start; k < NC; k += NT) {
            const int i = G.ri0 + r, j = G.rj0 + cc;
            float dz = S.dz[k];
            if (i >= 1 && j >= 1 && r >= 1 && cc >= 1) {
                const float curl = S.hy[k] - S.hy[k - RW] - S.hx[k] + S.hx[k - 1];
                dz = dz + 0.5f * curl;
                S.dz[k] = dz;
            }
            if (P.nbz) {
                const float iz = S.iz[k];
                const float ez = S.na[k] * (dz - iz);
                S.ez[k] = ez;
                S.iz[k] = iz + S.nb[k] * ez;
            } else {
                S.ez[k] = S.na[k] * dz;
            }
            r += DR; cc += DC;
            if (cc >= RW) { cc -= RW; ++r; }
        }
        __syncthreads();
        // ------------------------------------------------------ learned cells
        for (int k = tid; k < na * OUT; k += NT) {
            const int p = k / OUT, d = k % OUT;
            const int e = S.pos[p] >> 24, t = S.pos[p] & 0xFFFFFF;
            int i, j;
            cell_of(e, d, t, nx, ny, i, j);
            const int r = i - G.ri0, cc = j - G.rj0;
            if (r >= 0 && r < RH && cc >= 0 && cc < RW) S.ez[r * RW + cc] = S.pred[p][d];
        }
        __syncthreads();
        // ------------------------------------------------------ H update
        for (int k = tid, r = r_start, cc = c_start; k < NC; k += NT) {
            const int i = G.ri0 + r, j = G.rj0 + cc;
            if (i < nx - 1 && j < ny - 1 && r + 1 < RH && cc + 1 < RW) {
                const float ez = S.ez[k];
                S.hx[k] = S.hx[k] + 0.5f * (ez - S.ez[k + 1]);
                S.hy[k] = S.hy[k] - 0.5f * (ez - S.ez[k + RW]);
            }
            r += DR; cc += DC;
            if (cc >= RW) { cc -= RW; ++r; }
        }
        __syncthreads();
        // ------------------------------------------------------ new samples
        for (int k = tid; k < na * NS; k += NT) {
            const int p = k / NS, i = k % NS;
            const unsigned sv = S.src[p][i];
            const unsigned f = (sv >> 12) & 3u, cell = sv & 0xFFFu;
            float v = f == 0 ? S.ez[cell] : (f == 1 ? S.hx[cell] : (f == 2 ? S.hy[cell] : 0.0f));
            if (sv & 0x8000u) v = -v;
            S.ring[p][(-lev) & 3][i] = v;
        }
        __syncthreads();
    }

    // ---------------------------------------------------------- store
    for (int k = tid; k < NC; k += NT) {
        const int r = k / RW, cc = k % RW;
        const int i = G.ri0 + r, j = G.rj0 + cc;
        if (i < G.oi0 || i >= G.oi1 || j < G.oj0 || j >= G.oj1) continue;
        const long o = (long)i * ny + j;
        P.dz_o[o] = S.dz[k];
        if (P.iz_o) P.iz_o[o] = S.iz[k];
        P.hx_o[o] = S.hx[k];
        P.hy_o[o] = S.hy[k];
        if (P.ez_o) P.ez_o[o] = S.ez[k];
    }
    for (int k = tid; k < np * HIST * NS; k += NT) {
        const int p = k / (HIST * NS), a = (k / NS) % HIST, i = k % NS;
        if (S.own[p] < 0) continue;
        P.hist_out[((long)S.own[p] * HIST + a) * NS + i] = S.ring[p][(a - K) & 3][i];
    }
}

}  // namespace
}  // namespace lbc

int preload_lbc_frame()
{
    return preload_kernel(reinterpret_cast<const void *>(&lbc::frame_setup_kernel));
}

size_t lbc_frame_smem() { return sizeof(lbc::FrameSmem); }

size_t lbc_frame_table_bytes(int nx, int ny, int K)
{
    return (size_t)lbc_frame_blocks(nx, ny, K) * sizeof(lbc::FrameTab);
}

size_t lbc_frame_table_offset(int nx, int ny, int K)
{
    size_t o = 0;
    for (int k = 1; k < K; ++k) o += lbc_frame_table_bytes(nx, ny, k);
    return o;
}

int launch_lbc_frame_setup(int nx, int ny, int K, void *tab, hipStream_t s)
{
    hipLaunchKernelGGL(lbc::frame_setup_kernel, dim3(lbc_frame_blocks(nx, ny, K)), dim3(lbc::NT), 0,
                       s, nx, ny, K, reinterpret_cast<lbc::FrameTab *>(tab));
    return hip_check(hipGetLastError(), "lbc frame setup kernel");
}

int lbc_frame_blocks(int nx, int ny, int K)
{
    const int W = K + 3;
    return 2 * ((ny + lbc::CH - 1) / lbc::CH) + 2 * ((nx - 2 * W + lbc::CH - 1) / lbc::CH);
}

int launch_lbc_frame(const LbcFrameParams &p, hipStream_t s)
{
    static thread_local bool attr = false;
    if (!attr) {
        int rc = hip_check(hipFuncSetAttribute((const void *)lbc::frame_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)sizeof(lbc::FrameSmem)),
                           "lbc frame: shared memory");
        if (rc) return rc;
        attr = true;
    }
    hipLaunchKernelGGL(lbc::frame_kernel, dim3(lbc_frame_blocks(p.nx, p.ny, p.K)), dim3(lbc::NT),
                       sizeof(lbc::FrameSmem), s, p);
    return hip_check(hipGetLastError(), "lbc frame kernel");
}

}  // namespace fdtd
