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
// Each workgroup owns one 40-cell chunk of one edge's frame (depth K+3) and
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
// predicted cells replaced, H update, new samples pushed.  The ring enters
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

constexpr int CH = 40;                        // chunk length along the edge
constexpr int KMAX = 8;
// region cells: corner chunks (3K+6) x (CH + 2K + 1), others at most
// (2K+6) x (CH + 3K + 2) (30 x 57 and 22 x 66 at K = 8)
constexpr int RMAX_C = (3 * KMAX + 6) * (CH + 2 * KMAX + 1);
constexpr int RMAX_E = (2 * KMAX + 6) * (CH + 3 * KMAX + 2);
constexpr int RMAX = RMAX_C > RMAX_E ? RMAX_C : RMAX_E;
constexpr int PMAX = 128;                     // positions per workgroup

struct FrameSmem {
    float dz[RMAX], iz[RMAX], ez[RMAX], hx[RMAX], hy[RMAX], na[RMAX], nb[RMAX];
    float ring[PMAX][HIST][NS];
    uint16_t Xh[64 * XS];
    uint16_t A1[64 * AS], A2[64 * AS];
    float pred[PMAX][OUT];
    float srow[32];
    float lin[32][OUT];
    int pos[PMAX];          // edge << 24 | tau
    int own[PMAX];          // global history index, or -1 when not owned
    // where each sample value comes from: bits 0-11 region cell, 12-13 field
    // (0 ez, 1 hx, 2 hy, 3 constant 0), bit 15 negate
    uint16_t src[PMAX][NS];
};

static_assert(sizeof(FrameSmem) + 64 <= 160 * 1024, "frame kernel LDS above 160 KiB");

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

__global__ void __launch_bounds__(256) frame_kernel(LbcFrameParams P)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    FrameSmem &S = *reinterpret_cast<FrameSmem *>(smem_raw);
    const int nx = P.nx, ny = P.ny, K = P.K, W = K + 3;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, c = lane & 15, g = lane >> 4;
    const int nct = (ny + CH - 1) / CH;
    const int ncl = (nx - 2 * W + CH - 1) / CH;

    // ---------------------------------------------------------- geometry
    Geo G;
    int kind, q;                  // 0 top, 1 bottom, 2 left, 3 right
    {
        int b = blockIdx.x;
        if (b < nct) { kind = 0; q = b; }
        else if (b < 2 * nct) { kind = 1; q = b - nct; }
        else if (b < 2 * nct + ncl) { kind = 2; q = b - 2 * nct; }
        else { kind = 3; q = b - 2 * nct - ncl; }
    }
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
    const int RH = G.RH, RW = G.RW, NC = RH * RW;

    // ---------------------------------------------------------- positions
    // (thread 0 builds the list; at most CH + 3K + 2 + 2 (3K + 3) = 120)
    __shared__ int np_s;
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
        np_s = n;
    }

    // ---------------------------------------------------------- load
    for (int k = tid; k < RH * RW; k += 256) {
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
    __syncthreads();
    const int np = np_s;
    // history ring: sample of level l (l = -3..K) in slot l & 3
    for (int k = tid; k < np * HIST * NS; k += 256) {
        const int p = k / (HIST * NS), a = (k / NS) % HIST, i = k % NS;
        int e = S.pos[p] >> 24, t = S.pos[p] & 0xFFFFFF;
        S.ring[p][(-a) & 3][i] = P.hist_in[((long)gidx(e, t, nx, ny) * HIST + a) * NS + i];
    }
    __syncthreads();

    const Net N{P.w, P.params};
    // sample sources, resolved once: canon() run on an accessor that returns
    // the address instead of the value (clamped into the region: positions
    // near a cut side read stale neighbours, within the part of the region
    // that is no longer valid anyway)
    for (int k = tid; k < np * NS; k += 256) {
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
        S.src[p][i] = v;
    }
    // cell loops without divisions: thread tid starts at cell tid and steps
    // by 256 = DR rows + DC columns
    const int DR = 256 / RW, DC = 256 % RW;
    const int r_start = tid / RW, c_start = tid % RW;
    __syncthreads();

    for (int lev = 1; lev <= K; ++lev) {
        // ------------------------------------------------------ network
        // 32 positions per round: tile rows c and c + 8 hold position
        // base + 8 wave + (c & 7) and its negation (odd network)
        for (int base = 0; base < np; base += 32) {
            const int r0 = wave * 16;
            const int p = base + wave * 8 + (c & 7);
            float v[NF];
            if (p < np) {
                const float *src = S.ring[p][(lev - 1 - g) & 3];
#pragma unroll
                for (int i = 0; i < NF; ++i) v[i] = src[i];
            } else {
#pragma unroll
                for (int i = 0; i < NF; ++i) v[i] = 0.0f;
            }
            float s, inv;
            store_features(v, S.Xh, r0, s, inv);
            float lin[OUT];
            linear_out(P.params + P_L, v, inv, lin);
            if (g == 0 && c < 8) {
                S.srow[wave * 8 + c] = s;
#pragma unroll
                for (int d = 0; d < OUT; ++d) S.lin[wave * 8 + c][d] = lin[d];
            }
            __syncthreads();
            float z1[16], z2[16], y[4];
            forward_tile<false>(N, S.Xh, S.A1, S.A2, r0, z1, z2, y);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float yo = odd_out(y[k]);
                const int q8 = wave * 8 + g * 4 + k;          // valid for g < 2
                const int pp = base + q8;
                if (g < 2 && c < OUT && pp < np) {
                    const float sr = S.srow[q8];
                    const float b = base_of(c, S.ring[pp][(lev - 2) & 3]);
                    S.pred[pp][c] = sr > S_MIN ? b + sr * (S.lin[q8][c] + yo) : b;
                }
            }
            __syncthreads();
        }
        // ------------------------------------------------------ D, E update
        for (int k = tid, r = r_start, cc = c_start; k < NC; k += 256) {
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
        for (int k = tid; k < np * OUT; k += 256) {
            const int p = k / OUT, d = k % OUT;
            const int e = S.pos[p] >> 24, t = S.pos[p] & 0xFFFFFF;
            int i, j;
            cell_of(e, d, t, nx, ny, i, j);
            const int r = i - G.ri0, cc = j - G.rj0;
            if (r >= 0 && r < RH && cc >= 0 && cc < RW) S.ez[r * RW + cc] = S.pred[p][d];
        }
        __syncthreads();
        // ------------------------------------------------------ H update
        for (int k = tid, r = r_start, cc = c_start; k < NC; k += 256) {
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
        for (int k = tid; k < np * NS; k += 256) {
            const int p = k / NS, i = k % NS;
            const unsigned sv = S.src[p][i];
            const unsigned f = (sv >> 12) & 3u, cell = sv & 0xFFFu;
            float v = f == 0 ? S.ez[cell] : (f == 1 ? S.hx[cell] : (f == 2 ? S.hy[cell] : 0.0f));
            if (sv & 0x8000u) v = -v;
            S.ring[p][lev & 3][i] = v;
        }
        __syncthreads();
    }

    // ---------------------------------------------------------- store
    for (int k = tid; k < NC; k += 256) {
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
    for (int k = tid; k < np * HIST * NS; k += 256) {
        const int p = k / (HIST * NS), a = (k / NS) % HIST, i = k % NS;
        if (S.own[p] < 0) continue;
        P.hist_out[((long)S.own[p] * HIST + a) * NS + i] = S.ring[p][(K - a) & 3][i];
    }
}

}  // namespace
}  // namespace lbc

size_t lbc_frame_smem() { return sizeof(lbc::FrameSmem); }

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
    hipLaunchKernelGGL(lbc::frame_kernel, dim3(lbc_frame_blocks(p.nx, p.ny, p.K)), dim3(256),
                       sizeof(lbc::FrameSmem), s, p);
    return hip_check(hipGetLastError(), "lbc frame kernel");
}

}  // namespace fdtd
