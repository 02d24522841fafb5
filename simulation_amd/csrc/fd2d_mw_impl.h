// fd2d_mw_impl.h -- temporally blocked 2-D TM time stepping, multi-wave form
// (template body; instantiated per depth by fd2d_mw_k<depth>.hip).
//
// One launch advances the whole grid by K leapfrog steps, with the per-step
// semantics of the reference's kernels dfield -> hard source -> efield ->
// hfield (fd2d/pycuda/fd2d_3_2.py:219-260, fd2d/pycuda/fd2d_3_4.py:175-185),
// operand for operand, so the result equals K rounds of the per-step kernels
// bit for bit -- the same contract as tm_tb_kernel (fd2d_tb_impl.h), whose
// single-wave sweep this form extends.
//
// Why another form.  In tm_tb_kernel a wave of 64 lanes sweeps a strip alone:
// each fused step widens its dependency cone by one column per side, so only
// 64 - 2K of its columns are written (48 of 64 at K = 8), and short row blocks
// (for enough waves) add K warm-up and K drain rows per block.  Here a
// workgroup of W waves sweeps a strip of 64 W columns together, exchanging the
// columns at the wave boundaries through LDS, so the halo is paid only at the
// workgroup's two edges (240 of 256 columns written at W = 4, K = 8), and the
// strip's rows can be taller for the same number of workgroups.
//
// Level skew of two rows.  Level s (time step t_first + s) computes, at sweep
// position p, the E update of row r = p - 2s and the H update of row r - 1.
// With this skew everything a level takes from the neighbouring wave was
// produced one sweep position earlier:
//   * the E update of row r needs hx of row r after level s-1, column j-1:
//     level s-1 updated H row r at position p-1;
//   * the H update of row r-1 needs ez of row r-1 after level s, column j+1:
//     level s computed E row r-1 at position p-1.
// So per sweep position every wave reads its neighbours' edge values of the
// previous position from LDS, computes its K levels, writes its own edge
// values to the other LDS slot and meets the others at one barrier.  Within a
// position the levels do not depend on each other (they read only carries of
// earlier positions), and they run from K-1 down to 0 so that level s+1 reads
// the carries of level s before level s replaces them.
//
// Carries (registers, per level s < K-1, alternating by position parity):
//   D row of level s (dz, iz, naz, nbz)  produced at p, read by level s+1 at p+2
//   H row of level s (hx, hy, ihx, ihy)  produced at p, read by level s+1 at p+1
//                                        (as its E input) and p+2 (as the H
//                                        state its H update advances)
//   ez row of level s                    produced at p, read by level s at p+1
//
// Tiles.  The launch covers the grid with two kinds of workgroups:
//   * frame workgroups: the rows [0, RA) and [RB, nx), and the columns [0, CA)
//     and [CB, ny) of the rows between, as single-wave tiles of the general
//     body of fd2d_tb_impl.h (grid edges, PML profiles, any source or probe),
//     W tiles per workgroup, no barriers;
//   * interior workgroups: [RA, RB) x [CA, CB), whose sweeps stay inside the
//     band of default profile values, as multi-wave super-tiles (strips of
//     64 W - 2K written columns, `rows` rows), with sources and probes applied
//     where they fall.
// Frame workgroups come first (slower bodies start early); interior
// workgroups follow in an XCD-aware order (workgroup L runs on XCD L % 8; the
// workgroups of one XCD take a contiguous run of super-tiles, strips fastest).
#pragma once

#include "fd2d_tb_impl.h"

namespace fdtd {
namespace mw {

using tb::at;
using tb::DRow;
using tb::from_left;
using tb::Geom;
using tb::gfloat;
using tb::HRow;

constexpr int W = MW_WAVES;        // waves per workgroup
constexpr int NT = 64 * W;

// lane l receives lane l-1's value, lane 0 receives `fill`
__device__ __forceinline__ float shr1(float v, float fill)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                                 __builtin_bit_cast(int, v), 0x138,
                                                                 0xf, 0xf, false));
}
// lane l receives lane l+1's value, lane 63 receives `fill`
__device__ __forceinline__ float shl1(float v, float fill)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill),
                                                                 __builtin_bit_cast(int, v), 0x130,
                                                                 0xf, 0xf, false));
}

// LDS: the edge exchange, two slots by position parity.  Waves are stored at
// index w + 1; indices 0 and W + 1 stand for the missing neighbours of the
// workgroup's outer waves and stay zero (their columns are halo).
template <int K>
struct Xch {
    float hx[2][W + 2][K];        // lane 63's hx of each level's H row
    float ez[2][W + 2][K];        // lane 0's ez of each level's E row
    float src[FDTD_MAX_SOURCES][K];   // source values of the launch's K steps
};

// LDS: the coefficients naz, nbz of the last RING loaded rows of each wave
// (row r in slot r % RING).  Level s needs those of row p - 2s, loaded 2s
// positions earlier; a ring of 16 rows covers every depth up to 8.  Each
// wave writes and reads only its own ring, in order, so no barrier guards it.
constexpr int RING = 16;
struct alignas(16) CoefRing {
    float2 v[W][RING][64];
};

// The LDS wait and barrier between sweep positions.  Only LDS traffic must be
// complete (global loads in flight for the next row may stay in flight), so
// the wait is on the LDS counter only; the "memory" clobber keeps the
// compiler from moving LDS accesses across it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <bool LOSSY, bool PML>
__device__ __forceinline__ void load_row(const TbParams &P, unsigned row, unsigned col,
                                         HRow<1> &H, DRow<1> &D)
{
    const unsigned o = row + col;
    D.dz[0] = *at(P.dz_i, o);
    D.na[0] = *at(P.naz, o);
    H.hx[0] = *at(P.hx_i, o);
    H.hy[0] = *at(P.hy_i, o);
    if (LOSSY) {
        D.iz[0] = *at(P.iz_i, o);
        D.nb[0] = *at(P.nbz, o);
    } else {
        D.iz[0] = 0.f;
        D.nb[0] = 0.f;
    }
    if (PML) {
        H.ix[0] = *at(P.ihx_i, o);
        H.iy[0] = *at(P.ihy_i, o);
    } else {
        H.ix[0] = 0.f;
        H.iy[0] = 0.f;
    }
}

// Geometry of one interior super-tile sweep.
struct MwGeom {
    int i0, i1;        // written rows [i0, i1)
    int c0, c1;        // written columns [c0, c1)
    int j;             // the lane's column
    unsigned col;      // the lane's (clamped) load column
    int w, lane;
};

// One level of one sweep position.  E update of row r = p - 2s from the D
// row `Din` (the state after level s-1) and the H rows `Hin` (row r) and `Hp`
// (row r-1) after level s-1; H update of row r-1 from `Hp`, the ez row r-1 of
// this level (`Ec`, carried) and the new ez row r.  `hxl` / `ezr`: the
// neighbouring waves' edge values for lane 0 / lane 63.
template <int K, bool LOSSY, bool PML, bool SP, bool LAST>
__device__ __forceinline__ void level(const TbParams &P, const MwGeom &g, const Xch<K> &X, int s,
                                      int p, const DRow<1> &Din, float2 nab, const HRow<1> &Hin,
                                      const HRow<1> &Hp, float hxl, float ezr, float &Ec,
                                      DRow<1> &Dout, HRow<1> &Hout, float &ez_new)
{
    const int r = p - 2 * s;
    const int q = r - 1;
    // ------------------------------------------------ E update, row r
    const float hL = shr1(Hin.hx[0], hxl);
    const float curl = Hin.hy[0] - Hp.hy[0] - Hin.hx[0] + hL;
    // interior: the profile values are the defaults 1 (multiplications by
    // one are exact, so this is the general expression)
    float dz;
    if (PML) dz = 1.f * 1.f * Din.dz[0] + 1.f * 1.f * 0.5f * curl;
    else dz = Din.dz[0] + 0.5f * curl;
    if (SP) {
        // source values come from LDS (staged at the tile's start), so no
        // wait on the vector-memory counter lands in the sweep
        for (int k = 0; k < P.nsrc; ++k)
            if (r == P.src_i[k]) {
                const float v = X.src[k][s];
                dz = g.j == P.src_j[k] ? v : dz;
            }
    }
    float iz = Din.iz[0], ez;
    if (LOSSY) {
        ez = nab.x * (dz - iz);
        iz = iz + nab.y * ez;
    } else {
        ez = nab.x * dz;
    }
    if (SP && P.nprobe && r >= g.i0 && r < g.i1 && g.j >= g.c0 && g.j < g.c1) {
        const int t = P.t_first + s;
        for (int pr = 0; pr < P.nprobe; ++pr)
            if (r == P.probe_i[pr] && g.j == P.probe_j[pr])
                *(__attribute__((address_space(1))) float *)(P.probe_out + (size_t)pr * P.probe_len +
                                                             (t - 1)) = ez;
    }
    // ------------------------------------------------ H update, row q
    const float e = Ec;
    const float eR = shl1(Ec, ezr);
    const float eD = ez;
    float hx, hy, ax, ay;
    if (PML) {
        ax = Hp.ix[0] + (e - eR);
        ay = Hp.iy[0] + (e - eD);
        hx = 1.f * Hp.hx[0] + 1.f * (0.5f * e - 0.5f * eR + 0.f * ax);
        hy = 1.f * Hp.hy[0] - 1.f * (0.5f * e - 0.5f * eD + 0.f * ay);
    } else {
        ax = Hp.ix[0];
        ay = Hp.iy[0];
        hx = Hp.hx[0] + 0.5f * (e - eR);
        hy = Hp.hy[0] - 0.5f * (e - eD);
    }
    Ec = ez;
    ez_new = ez;
    if (LAST) {
        // final values of this launch: D/E of row r and H of row q
        const unsigned nb = (unsigned)P.nx * (unsigned)P.ny * 4u;
        const bool okc = g.j >= g.c0 && g.j < g.c1;
        const unsigned jr = (unsigned)g.j;
        const bool okr = okc && r >= g.i0 && r < g.i1;
        const unsigned offr = okr ? ((unsigned)r * (unsigned)P.ny + jr) << 2 : tb::OOB;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, dz), tb::out_rsrc(P.dz_o, nb),
                                              offr, 0, 0);
        if (LOSSY)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, iz),
                                                  tb::out_rsrc(P.iz_o, nb), offr, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, ez),
                                              tb::out_rsrc(P.ez_o, P.ez_o ? nb : 0u), offr, 0, 0);
        const bool okq = okc && q >= g.i0 && q < g.i1;
        const unsigned offq = okq ? ((unsigned)q * (unsigned)P.ny + jr) << 2 : tb::OOB;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, hx), tb::out_rsrc(P.hx_o, nb),
                                              offq, 0, 0);
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, hy), tb::out_rsrc(P.hy_o, nb),
                                              offq, 0, 0);
        if (PML) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, ax),
                                                  tb::out_rsrc(P.ihx_o, nb), offq, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, ay),
                                                  tb::out_rsrc(P.ihy_o, nb), offq, 0, 0);
        }
    } else {
        Dout.dz[0] = dz;
        Dout.iz[0] = iz;
        Hout.hx[0] = hx;
        Hout.hy[0] = hy;
        Hout.ix[0] = ax;
        Hout.iy[0] = ay;
    }
}

__device__ __forceinline__ float uniform(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// Level s is active at positions [beg(s), end(s)): E rows [i0-K+1+s,
// i1+K-s) and H rows [i0-K+1+s, i1+K-1-s) -- the rows the written rows of
// the last level depend on.  The H update of the first active position
// updates a row no later level reads; it is computed with the rest.
template <int K>
__device__ __forceinline__ int lv_beg(int i0, int s) { return i0 - K + 1 + 3 * s; }
template <int K>
__device__ __forceinline__ int lv_end(int i1, int s) { return i1 + K + s; }

// The sweep of one super-tile by the W waves of a workgroup.
template <int K, bool LOSSY, bool PML, bool SP>
__device__ __forceinline__ void sweep(const TbParams &P, const MwGeom &g, Xch<K> &X, CoefRing &R)
{
    // carries: [level][parity]; the D carries hold dz and iz only (naz and
    // nbz come from the coefficient ring)
    DRow<1> Dc[K > 1 ? K - 1 : 1][2];
    HRow<1> Hc[K > 1 ? K - 1 : 1][2];
    float Ec[K];
#pragma unroll
    for (int s = 0; s < K; ++s) Ec[s] = 0.f;
#pragma unroll
    for (int s = 0; s + 1 < K; ++s)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            Dc[s][b].dz[0] = Dc[s][b].iz[0] = Dc[s][b].na[0] = Dc[s][b].nb[0] = 0.f;
            Hc[s][b].hx[0] = Hc[s][b].hy[0] = Hc[s][b].ix[0] = Hc[s][b].iy[0] = 0.f;
        }
    const int pbeg = g.i0 - K + 1;              // first position (level 0)
    // one past the last position (level K-1), rounded up to whole loop
    // trips: trailing positions with no active level only load and meet
    // the barrier
    constexpr int U = 6;                         // positions per loop trip
    const int pend = pbeg + (g.i1 + 2 * K - 1 - pbeg + U - 1) / U * U;
    const int rlast = g.i1 + K - 1;              // last row level 0 reads
    const unsigned ny = (unsigned)P.ny;
    float2 (&ring)[RING][64] = R.v[g.w];
    // the column left of the wave's first, for lane 0's E update at level
    // 0: a wave-uniform address, loaded by every lane (one transaction, no
    // branch around the load)
    const unsigned xcol = (unsigned)min(g.j - g.lane - 1, P.ny - 1);
    // Loaded rows.  Position p uses the H rows p (E input of level 0) and
    // p-1 (its H state) and prefetches row p+1: three H buffers in turn
    // (position p: buffer p % 3), two D buffers and two left-column values
    // (p % 2), so no loaded register is ever copied -- a copy would have to
    // wait for the load in flight.  With the carry parity that makes six
    // positions per loop trip, each with compile-time buffer indices.
    HRow<1> Hb[3];
    DRow<1> Db[2];
    float xb[2];
    load_row<LOSSY, PML>(P, (unsigned)(pbeg - 1) * ny, g.col, Hb[2], Db[1]);
    load_row<LOSSY, PML>(P, (unsigned)pbeg * ny, g.col, Hb[0], Db[0]);
    xb[0] = *at(P.hx_i, (unsigned)pbeg * ny + xcol);
#pragma unroll 1
    for (int p = pbeg; p < pend; p += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int pp = p + u;
            const int b = u & 1;                 // parity of pp relative to pbeg
            HRow<1> &Hcur = Hb[u % 3];
            HRow<1> &Hprev = Hb[(u + 2) % 3];
            HRow<1> &Hnxt = Hb[(u + 1) % 3];
            DRow<1> &Dcur = Db[b];
            DRow<1> &Dnxt = Db[b ^ 1];
            const float xcur = xb[b];
            // prefetch the row of the next position (rows past the last
            // loaded row are never used; they read the last row again)
            {
                const int rn = min(pp + 1, rlast);
                load_row<LOSSY, PML>(P, (unsigned)rn * ny, g.col, Hnxt, Dnxt);
                xb[b ^ 1] = *at(P.hx_i, (unsigned)rn * ny + xcol);
            }
            const int sl = b ^ 1;                // LDS slot of the previous position
            // neighbours' edge values of the previous position, wave-uniform
            // (held in scalar registers): the left wave's hx of each level
            // (index w) and the right wave's ez of each level (index w+2)
            float hxl[K], ezr[K];
#pragma unroll
            for (int s = 0; s < K; ++s) {
                hxl[s] = uniform(X.hx[sl][g.w][s]);
                ezr[s] = uniform(X.ez[sl][g.w + 2][s]);
            }
            float hxo[K], ezo[K];
            // the coefficients of level s's row come from the ring, read one
            // level ahead of their use
            float2 nab_next = make_float2(0.f, 0.f);
            if (K > 1) nab_next = ring[(pp - 2 * (K - 1)) & (RING - 1)][g.lane];
#pragma unroll
            for (int s = K - 1; s >= 0; --s) {
                hxo[s] = 0.f;
                ezo[s] = 0.f;
                const float2 nab_s = nab_next;
                if (s >= 2) nab_next = ring[(pp - 2 * (s - 1)) & (RING - 1)][g.lane];
                const bool act = pp >= lv_beg<K>(g.i0, s) && pp < lv_end<K>(g.i1, s);
                if (!act) continue;
                if (s == 0) {
                    const float2 nab = make_float2(Dcur.na[0], Dcur.nb[0]);
                    if (K == 1) {
                        DRow<1> dd; HRow<1> hh;
                        level<K, LOSSY, PML, SP, true>(P, g, X, 0, pp, Dcur, nab, Hcur, Hprev, xcur,
                                                       ezr[0], Ec[0], dd, hh, ezo[0]);
                    } else {
                        level<K, LOSSY, PML, SP, false>(P, g, X, 0, pp, Dcur, nab, Hcur, Hprev, xcur,
                                                        ezr[0], Ec[0], Dc[0][b], Hc[0][b], ezo[0]);
                        hxo[0] = Hc[0][b].hx[0];
                    }
                } else {
                    const float2 nab = nab_s;
                    if (s == K - 1) {
                        DRow<1> dd; HRow<1> hh;
                        level<K, LOSSY, PML, SP, true>(P, g, X, s, pp, Dc[s - 1][b], nab,
                                                       Hc[s - 1][b ^ 1], Hc[s - 1][b], hxl[s - 1],
                                                       ezr[s], Ec[s], dd, hh, ezo[s]);
                    } else {
                        level<K, LOSSY, PML, SP, false>(P, g, X, s, pp, Dc[s - 1][b], nab,
                                                        Hc[s - 1][b ^ 1], Hc[s - 1][b], hxl[s - 1],
                                                        ezr[s], Ec[s], Dc[s][b], Hc[s][b], ezo[s]);
                        hxo[s] = Hc[s][b].hx[0];
                    }
                }
            }
            // the coefficients of the row level 0 used, into the ring (after
            // level 0, which waits for the row anyway; levels s >= 1 read
            // rows at least two positions old)
            ring[pp & (RING - 1)][g.lane] = make_float2(Dcur.na[0], Dcur.nb[0]);
            // this position's edge values for the neighbours
            if (g.lane == 63) {
#pragma unroll
                for (int s = 0; s < K; ++s) X.hx[b][g.w + 1][s] = hxo[s];
            }
            if (g.lane == 0) {
#pragma unroll
                for (int s = 0; s < K; ++s) X.ez[b][g.w + 1][s] = ezo[s];
            }
            lds_barrier();
        }
    }
}

// The launch's geometry: the frame bands (single-wave tiles of the general
// body) and the interior super-tiles.
struct MwParams {
    TbParams P;
    int ra, rb, ca, cb;         // interior rectangle [ra, rb) x [ca, cb)
    int rows;                   // rows of an interior super-tile
    int nsx, nrb;               // interior super-strips and row blocks
    int frows;                  // rows of a frame tile
    int band[4][6];             // frame bands: r0, r1, c0, c1, strips, row blocks
    int ntf;                    // frame tiles
    int nfw;                    // frame workgroups (padded to a multiple of 8)
};

template <int K, bool LOSSY, bool PML>
__global__ void __launch_bounds__(NT) tm_mw_kernel(const MwParams M)
{
    // the interior workgroups' exchange buffers and coefficient rings, or the
    // frame workgroups' single-wave profile rings
    __shared__ union {
        struct {
            Xch<K> X;
            CoefRing R;
        } in;
        tb::NaRing<K, 1> fr[W];
    } U;
    Xch<K> &X = U.in.X;
    CoefRing &R = U.in.R;
    const TbParams &P = M.P;
    const int w = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
    if ((int)blockIdx.x < M.nfw) {
        // frame: W single-wave tiles per workgroup, the general body
        int t = (int)blockIdx.x * W + w;
        if (t >= M.ntf) return;
        int b = 0;
        while (t >= M.band[b][4] * M.band[b][5]) {
            t -= M.band[b][4] * M.band[b][5];
            ++b;
        }
        const int fw = 64 - 2 * P.halo;
        const int x = t % M.band[b][4], y = t / M.band[b][4];
        Geom g;
        g.vj0 = M.band[b][2] + x * fw;
        g.vj1 = min(M.band[b][3], g.vj0 + fw);
        g.j0 = g.vj0 - P.halo + lane;
        g.i0 = M.band[b][0] + y * M.frows;
        g.i1 = min(M.band[b][1], g.i0 + M.frows);
        tb::sweep<K, 1, LOSSY, PML, false, true, true, true, 0, true>(P, g, U.fr[w]);
        return;
    }
    // interior super-tile, XCD-aware order
    const unsigned L = blockIdx.x - (unsigned)M.nfw, nt = gridDim.x - (unsigned)M.nfw;
    const unsigned xcd = L % 8u, qd = nt / 8u, rem = nt % 8u;
    const unsigned tile = xcd * qd + min(xcd, rem) + L / 8u;
    const int bx = (int)(tile % (unsigned)M.nsx), by = (int)(tile / (unsigned)M.nsx);
    for (int k = (int)threadIdx.x; k < (int)(sizeof(Xch<K>) / sizeof(float)); k += NT)
        reinterpret_cast<float *>(&X)[k] = 0.f;
    __syncthreads();
    for (int k = (int)threadIdx.x; k < P.nsrc * K; k += NT)
        X.src[k / K][k % K] = *(gfloat *)(P.src_wave + (size_t)(k / K) * P.wave_len +
                                          (P.t_first + k % K - 1));
    __syncthreads();
    constexpr int SW = NT - 2 * K;
    MwGeom g;
    g.i0 = M.ra + by * M.rows;
    g.i1 = min(M.rb, g.i0 + M.rows);
    g.c0 = M.ca + bx * SW;
    g.c1 = min(M.cb, g.c0 + SW);
    const int js = g.c0 - K;
    g.j = js + w * 64 + lane;
    g.col = (unsigned)min(g.j, P.ny - 1);
    g.w = w;
    g.lane = lane;
    // a source in the rows and columns the sweep computes, or a probe in the
    // written cells (workgroup-uniform)
    bool sp = false;
    for (int k = 0; k < P.nsrc; ++k)
        sp |= P.src_i[k] >= g.i0 - K && P.src_i[k] < g.i1 + K && P.src_j[k] >= js &&
              P.src_j[k] < js + NT;
    for (int k = 0; k < P.nprobe; ++k)
        sp |= P.probe_i[k] >= g.i0 && P.probe_i[k] < g.i1 && P.probe_j[k] >= g.c0 &&
              P.probe_j[k] < g.c1;
    if (sp) sweep<K, LOSSY, PML, true>(P, g, X, R);
    else sweep<K, LOSSY, PML, false>(P, g, X, R);
}

// Frame and interior of a launch.  The interior sweeps read rows [i0-K,
// i1+K) and columns [c0-K-1, c1+K) of their super-tiles; those must lie in
// the band of default profile values [rlo, rhi) x [clo, chi) (itself inside
// [1, n-2]).
inline void mw_geometry(MwParams &m, int K, int rows, int frows)
{
    const TbParams &p = m.P;
    m.ra = p.rlo + K;
    m.rb = p.rhi - K;
    m.ca = p.clo + K + 1;
    m.cb = p.chi - K;
    if (m.rb - m.ra < 1 || m.cb - m.ca < 1 || p.rhi <= p.rlo || p.chi <= p.clo) {
        m.ra = m.rb = p.nx;       // no interior: the frame is the whole grid
        m.ca = m.cb = 0;
    }
    m.rows = rows;
    m.frows = frows;
    const int SW = NT - 2 * K;
    m.nsx = m.rb > m.ra ? (m.cb - m.ca + SW - 1) / SW : 0;
    m.nrb = m.rb > m.ra ? (m.rb - m.ra + rows - 1) / rows : 0;
    const int fw = 64 - 2 * p.halo;
    const int bands[4][4] = {{0, m.ra, 0, p.ny}, {m.rb, p.nx, 0, p.ny},
                             {m.ra, m.rb, 0, m.ca}, {m.ra, m.rb, m.cb, p.ny}};
    m.ntf = 0;
    for (int b = 0; b < 4; ++b) {
        const int r0 = bands[b][0], r1 = bands[b][1], c0 = bands[b][2], c1 = bands[b][3];
        const bool empty = r1 <= r0 || c1 <= c0;
        m.band[b][0] = r0; m.band[b][1] = r1; m.band[b][2] = c0; m.band[b][3] = c1;
        m.band[b][4] = empty ? 0 : (c1 - c0 + fw - 1) / fw;
        m.band[b][5] = empty ? 0 : (r1 - r0 + frows - 1) / frows;
        m.ntf += m.band[b][4] * m.band[b][5];
    }
    m.nfw = (m.ntf + W - 1) / W;
    m.nfw = (m.nfw + 7) / 8 * 8;
}

// Host check of the bounds the interior sweeps rely on: rows [ra-K, rb+K)
// and columns [ca-K-1, cb+K) inside the grid, frame tiles no wider than a
// wave's strip.
inline bool mw_geometry_ok(const MwParams &m, int K)
{
    const TbParams &p = m.P;
    if (m.rows < 1 || m.frows < 1 || p.halo < K || 64 - 2 * p.halo < 8) return false;
    if (m.rb > m.ra && (m.ra - K < 0 || m.rb + K > p.nx || m.ca - K - 1 < 0 || m.cb + K > p.ny))
        return false;
    return true;
}

template <int K>
int launch_mw(const TbParams &p, bool lossy, bool pml, int rows, int frows, hipStream_t s)
{
    MwParams m;
    m.P = p;
    m.P.n_rb = -1;
    m.P.n_sx = m.P.n_t = 0;
    m.P.nb = m.P.nb_pad = 0;
    mw_geometry(m, K, rows, frows);
    if (!mw_geometry_ok(m, K))
        return fail("fd2d_tm_run: multi-wave geometry (rows %d, frame rows %d, halo %d) rejected",
                    rows, frows, p.halo);
    const long nwg = (long)m.nfw + (long)m.nsx * (long)m.nrb;
    if (nwg <= 0) return 0;
    if (nwg > 0x7fffffffL) return fail("fd2d_tm_run: %ld workgroups", nwg);
    const dim3 grid((unsigned)nwg), block(NT);
    if (lossy && pml) hipLaunchKernelGGL((tm_mw_kernel<K, true, true>), grid, block, 0, s, m);
    else if (lossy) hipLaunchKernelGGL((tm_mw_kernel<K, true, false>), grid, block, 0, s, m);
    else if (pml) hipLaunchKernelGGL((tm_mw_kernel<K, false, true>), grid, block, 0, s, m);
    else hipLaunchKernelGGL((tm_mw_kernel<K, false, false>), grid, block, 0, s, m);
    return hip_check(hipGetLastError(), "fd2d_tm_run launch (multi-wave)");
}

}  // namespace mw

template <int K>
int launch_fd2d_mw_depth(const TbParams &p, bool lossy, bool pml, int rows, int frows, hipStream_t s)
{
    return mw::launch_mw<K>(p, lossy, pml, rows, frows, s);
}

template <int K>
int preload_mw_depth()
{
    return preload_kernel(reinterpret_cast<const void *>(&mw::tm_mw_kernel<K, true, true>));
}

}  // namespace fdtd
