// fd2d_tb.hip -- fused, temporally blocked 2-D TM time stepping for gfx950.
//
// One launch advances the whole grid by K leapfrog steps (K = template
// depth).  Semantics per step are those of the reference's per-step kernels
// dfield -> hard source -> efield -> hfield (fd2d/pycuda/fd2d_3_2.py:219-260,
// fd2d/pycuda/fd2d_3_4.py:175-185), operand for operand, so the result equals
// K rounds of fd2d_dfield/efield/hfield bit for bit.
//
// Decomposition (one wavefront per workgroup, no barriers):
//   * Columns: a wave owns a strip of S = 64*C columns, lane l holding the C
//     contiguous columns js + l*C .. js + l*C + C-1.  Neighbour columns come
//     from the same lane or from lane l-1 / l+1 by a cross-lane shuffle.
//     Each fused step widens the dependency cone by one column on each side,
//     so a strip writes only its central S - 2*halo columns; neighbouring
//     strips overlap by the halo (recomputed, never written twice).
//   * Rows: the wave sweeps down a block of `rows` rows.  Time level s of the
//     K fused steps trails level s-1 by one row: at sweep position p, level s
//     computes the E update of row p-s+1 and the H update of row p-s, fed by
//     the H row that level s-1 produced in the same iteration and the D row it
//     produced one iteration earlier.  Everything a level needs from the
//     previous row is carried in registers, so HBM is read once (level 1,
//     row p) and written once (level K) per launch.  A block starts K rows
//     above its first row (warm-up) and ends K rows below its last (drain).
//   * Launches read one state set and write the other (ping-pong), so waves
//     never read a cell another wave has already advanced.
//
// HBM traffic per cell per launch (lossy PML form): reads dz iz hx hy ihx ihy
// naz nbz (32 B, x halo/warm-up overlap), writes dz iz hx hy ihx ihy (24 B,
// + 4 B ez on the last launch of a run), for K cell-updates.
#include "fdtd_internal.h"

namespace fdtd {
namespace {

constexpr int kMaxDepth = 16;

__device__ __forceinline__ float from_left(float v) { return __shfl_up(v, 1, 64); }
__device__ __forceinline__ float from_right(float v) { return __shfl_down(v, 1, 64); }

template <int C>
struct HRow { float hx[C], hy[C], ix[C], iy[C]; };
template <int C>
struct DRow { float dz[C], iz[C], na[C], nb[C]; };

template <int C, bool LOSSY, bool PML>
struct Cols {
    int j[C];
    bool in[C], dupd[C], hupd[C], own[C];
    float gy2[C], gy3[C], fy1[C], fy2[C], fy3[C];
};

template <int C, bool LOSSY, bool PML>
__device__ __forceinline__ void load_row(const TbParams &P, const Cols<C, LOSSY, PML> &cl,
                                         int p, HRow<C> &H, DRow<C> &D)
{
#pragma unroll
    for (int c = 0; c < C; ++c) {
        H.hx[c] = 0.f; H.hy[c] = 0.f; H.ix[c] = 0.f; H.iy[c] = 0.f;
        D.dz[c] = 0.f; D.iz[c] = 0.f; D.na[c] = 0.f; D.nb[c] = 0.f;
    }
    if (p < 0 || p >= P.nx) return;
    const size_t base = (size_t)p * P.ny;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if (!cl.in[c]) continue;
        const size_t n = base + cl.j[c];
        D.dz[c] = P.dz_i[n];
        D.na[c] = P.naz[n];
        H.hx[c] = P.hx_i[n];
        H.hy[c] = P.hy_i[n];
        if (LOSSY) { D.iz[c] = P.iz_i[n]; D.nb[c] = P.nbz[n]; }
        if (PML) { H.ix[c] = P.ihx_i[n]; H.iy[c] = P.ihy_i[n]; }
    }
}

// One time level at sweep position p: E update of row r, H update of row r-1.
//   Hin  : H of level s-1 at row r (this iteration)
//   Din  : D/E state of level s-1 at row r (with that row's naz, nbz)
//   Hp   : H of level s-1 at row r-1 (previous Hin), updated to Hin on exit
//   Ep   : ez of this level at row r-1, updated to this row's ez on exit
//   EpR  : ez of this level at row r-1, column to the right of the lane's last
//   Hout : H of this level at row r-1 (input of level s+1)
//   Dout : D/E state of this level at row r
template <int C, bool LOSSY, bool PML>
__device__ __forceinline__ void level(const TbParams &P, const Cols<C, LOSSY, PML> &cl,
                                      int r, int t, const HRow<C> &Hin, const DRow<C> &Din,
                                      HRow<C> &Hp, float (&Ep)[C], float &EpR,
                                      HRow<C> &Hout, DRow<C> &Dout, bool owned_r)
{
    const int nx = P.nx;
    const int q = r - 1;
    float En[C];
    // ---- E update of row r
    if (r >= 0 && r < nx) {
        const float hxl = from_left(Hin.hx[C - 1]);
        float gx2r = 1.f, gx3r = 1.f;
        if (PML && r >= 1) { gx2r = P.pml.gx2[r]; gx3r = P.pml.gx3[r]; }
        const bool rowd = r >= 1;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float hL = (c == 0) ? hxl : Hin.hx[c - 1];
            float dz = Din.dz[c];
            if (rowd && cl.dupd[c]) {
                const float curl = Hin.hy[c] - Hp.hy[c] - Hin.hx[c] + hL;
                if (PML) dz = gx3r * cl.gy3[c] * dz + gx2r * cl.gy2[c] * 0.5f * curl;
                else dz = dz + 0.5f * curl;
            }
            for (int k = 0; k < P.nsrc; ++k)
                if (r == P.src_i[k] && cl.j[c] == P.src_j[k])
                    dz = P.src_wave[(size_t)k * P.wave_len + (t - 1)];
            float iz = Din.iz[c], ez;
            if (LOSSY) {
                ez = Din.na[c] * (dz - iz);
                iz = iz + Din.nb[c] * ez;
            } else {
                ez = Din.na[c] * dz;
            }
            Dout.dz[c] = dz; Dout.iz[c] = iz; Dout.na[c] = Din.na[c]; Dout.nb[c] = Din.nb[c];
            En[c] = ez;
        }
        if (owned_r) {
            for (int pr = 0; pr < P.nprobe; ++pr)
                if (r == P.probe_i[pr]) {
#pragma unroll
                    for (int c = 0; c < C; ++c)
                        if (cl.own[c] && cl.j[c] == P.probe_j[pr])
                            P.probe_out[(size_t)pr * P.probe_len + (t - 1)] = En[c];
                }
        }
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            En[c] = 0.f;
            Dout.dz[c] = 0.f; Dout.iz[c] = 0.f; Dout.na[c] = 0.f; Dout.nb[c] = 0.f;
        }
    }
    // ---- H update of row q = r-1
    if (q >= 0 && q < nx) {
        const bool rowh = q <= nx - 2;
        float fx1q = 0.f, fx2q = 1.f, fx3q = 1.f;
        if (PML && rowh) { fx1q = P.pml.fx1[q]; fx2q = P.pml.fx2[q]; fx3q = P.pml.fx3[q]; }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float e = Ep[c];
            const float eR = (c + 1 < C) ? Ep[c + 1] : EpR;
            const float eD = En[c];
            if (rowh && cl.hupd[c]) {
                if (PML) {
                    const float ax = Hp.ix[c] + (e - eR);
                    const float ay = Hp.iy[c] + (e - eD);
                    Hout.hx[c] = cl.fy3[c] * Hp.hx[c] + cl.fy2[c] * (0.5f * e - 0.5f * eR + fx1q * ax);
                    Hout.hy[c] = fx3q * Hp.hy[c] - fx2q * (0.5f * e - 0.5f * eD + cl.fy1[c] * ay);
                    Hout.ix[c] = ax;
                    Hout.iy[c] = ay;
                } else {
                    Hout.hx[c] = Hp.hx[c] + 0.5f * (e - eR);
                    Hout.hy[c] = Hp.hy[c] - 0.5f * (e - eD);
                    Hout.ix[c] = Hp.ix[c];
                    Hout.iy[c] = Hp.iy[c];
                }
            } else {
                Hout.hx[c] = Hp.hx[c]; Hout.hy[c] = Hp.hy[c];
                Hout.ix[c] = Hp.ix[c]; Hout.iy[c] = Hp.iy[c];
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c) { Hout.hx[c] = 0.f; Hout.hy[c] = 0.f; Hout.ix[c] = 0.f; Hout.iy[c] = 0.f; }
    }
    // ---- carry this row to the next iteration
    Hp = Hin;
#pragma unroll
    for (int c = 0; c < C; ++c) Ep[c] = En[c];
    EpR = from_right(En[0]);
}

template <int K, int C, bool LOSSY, bool PML>
__global__ void __launch_bounds__(64) tm_tb_kernel(const TbParams P)
{
    constexpr int S = 64 * C;
    const int lane = threadIdx.x;
    const int nx = P.nx, ny = P.ny;
    const int js = (int)blockIdx.x * (S - 2 * P.halo) - P.halo;
    const int i0 = (int)blockIdx.y * P.rows;
    const int i1 = min(i0 + P.rows, nx);
    const int vj0 = js + P.halo;
    const int vj1 = min(js + S - P.halo, ny);

    Cols<C, LOSSY, PML> cl;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = js + lane * C + c;
        cl.j[c] = j;
        cl.in[c] = (j >= 0) && (j < ny);
        cl.dupd[c] = (j >= 1) && (j < ny);
        cl.hupd[c] = (j >= 0) && (j <= ny - 2);
        cl.own[c] = (j >= vj0) && (j < vj1);
        cl.gy2[c] = 1.f; cl.gy3[c] = 1.f; cl.fy1[c] = 0.f; cl.fy2[c] = 1.f; cl.fy3[c] = 1.f;
        if (PML && cl.in[c]) {
            cl.gy2[c] = P.pml.gy2[j]; cl.gy3[c] = P.pml.gy3[j];
            cl.fy1[c] = P.pml.fy1[j]; cl.fy2[c] = P.pml.fy2[j]; cl.fy3[c] = P.pml.fy3[j];
        }
    }

    HRow<C> Hp[K];
    float Ep[K][C];
    float EpR[K];
    DRow<C> Q[K];
#pragma unroll
    for (int s = 0; s < K; ++s) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            Hp[s].hx[c] = 0.f; Hp[s].hy[c] = 0.f; Hp[s].ix[c] = 0.f; Hp[s].iy[c] = 0.f;
            Ep[s][c] = 0.f;
            Q[s].dz[c] = 0.f; Q[s].iz[c] = 0.f; Q[s].na[c] = 0.f; Q[s].nb[c] = 0.f;
        }
        EpR[s] = 0.f;
    }

    const int pbeg = max(i0 - K, 0);
    const int pend = i1 - 1 + K;
    HRow<C> Hn;
    DRow<C> Dn;
    load_row(P, cl, pbeg, Hn, Dn);

    for (int p = pbeg; p <= pend; ++p) {
        HRow<C> Hin = Hn;
        DRow<C> Din = Dn;
        if (p + 1 <= pend) load_row(P, cl, p + 1, Hn, Dn);   // prefetch the next row

        DRow<C> pend_d;      // D output of the previous level in this iteration
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int r = p - s;                 // E row of level s+1
            const int t = P.t_first + s;         // its time index
            DRow<C> din;
            if (s == 0) {
                din = Din;
            } else {
                din = Q[s - 1];                  // level s's D output from the last iteration
                Q[s - 1] = pend_d;               // ... replaced by this iteration's
            }
            HRow<C> Hout;
            DRow<C> Dout;
            const bool owned_r = (r >= i0) && (r < i1);
            level<C, LOSSY, PML>(P, cl, r, t, Hin, din, Hp[s], Ep[s], EpR[s], Hout, Dout, owned_r);
            pend_d = Dout;
            Hin = Hout;
            if (s == K - 1) {
                // level K: D/E row r and H row r-1 are final for this launch
                if (owned_r) {
                    const size_t base = (size_t)r * ny;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        if (!cl.own[c]) continue;
                        const size_t n = base + cl.j[c];
                        P.dz_o[n] = Dout.dz[c];
                        if (LOSSY) P.iz_o[n] = Dout.iz[c];
                        if (P.ez_o) P.ez_o[n] = Ep[s][c];
                    }
                }
                const int q = r - 1;
                if (q >= i0 && q < i1) {
                    const size_t base = (size_t)q * ny;
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        if (!cl.own[c]) continue;
                        const size_t n = base + cl.j[c];
                        P.hx_o[n] = Hout.hx[c];
                        P.hy_o[n] = Hout.hy[c];
                        if (PML) { P.ihx_o[n] = Hout.ix[c]; P.ihy_o[n] = Hout.iy[c]; }
                    }
                }
            }
        }
    }
}

template <int K, int C>
int launch_k(const TbParams &p, bool lossy, bool pml, hipStream_t s)
{
    constexpr int S = 64 * C;
    const int w = S - 2 * p.halo;
    const dim3 grid((p.ny + w - 1) / w, (p.nx + p.rows - 1) / p.rows);
    if (lossy && pml) hipLaunchKernelGGL((tm_tb_kernel<K, C, true, true>), grid, dim3(64), 0, s, p);
    else if (lossy) hipLaunchKernelGGL((tm_tb_kernel<K, C, true, false>), grid, dim3(64), 0, s, p);
    else if (pml) hipLaunchKernelGGL((tm_tb_kernel<K, C, false, true>), grid, dim3(64), 0, s, p);
    else hipLaunchKernelGGL((tm_tb_kernel<K, C, false, false>), grid, dim3(64), 0, s, p);
    return hip_check(hipGetLastError(), "fd2d_tm_run launch");
}

template <int C>
int launch_c(const TbParams &p, int depth, bool lossy, bool pml, hipStream_t s)
{
    switch (depth) {
    case 1: return launch_k<1, C>(p, lossy, pml, s);
    case 2: return launch_k<2, C>(p, lossy, pml, s);
    case 3: return launch_k<3, C>(p, lossy, pml, s);
    case 4: return launch_k<4, C>(p, lossy, pml, s);
    case 5: return launch_k<5, C>(p, lossy, pml, s);
    case 6: return launch_k<6, C>(p, lossy, pml, s);
    case 7: return launch_k<7, C>(p, lossy, pml, s);
    case 8: return launch_k<8, C>(p, lossy, pml, s);
    default: return fail("temporal depth %d not built", depth);
    }
}

}  // namespace

int tb_max_depth() { return 8; }

bool tb_supported(int depth, int cols)
{
    return depth >= 1 && depth <= tb_max_depth() && (cols == 1 || cols == 2);
}

int launch_fd2d_tb(const TbParams &p, int depth, int cols, bool lossy, bool pml, hipStream_t s)
{
    if (cols == 1) return launch_c<1>(p, depth, lossy, pml, s);
    if (cols == 2) return launch_c<2>(p, depth, lossy, pml, s);
    return fail("columns per lane %d not built", cols);
}

}  // namespace fdtd
