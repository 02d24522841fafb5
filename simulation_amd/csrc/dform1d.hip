// 1-D D/E form with lossy and Debye media and the 1-D running DFT
// (fd1d/pycuda/fd1d_2_1.py, fd1d_2_2.py, fd1d_2_3.py:73-158; serial semantics
// of fd1d/clang/test_2_3.c:26-90).
//
// Per step: D update with an additive source at one cell, E from D through
// the medium (nax, nbx; with ncx, ndx the Debye relaxation sx), the running
// DFT of ex, then the two-step absorbing boundary and the H update (the
// existing fd1d_hyfield).  FP32 in the reference's operand order without
// contraction; the DFT twiddles are evaluated on the host.
#include <hip/hip_runtime.h>
#include <math.h>

#include "fdtd_internal.h"

namespace fdtd {
namespace {

// dx[i] += 0.5 (hy[i-1] - hy[i]) for i in [1, nx); the thread that owns
// src_pos then adds the source (after its own update, as the serial loop).
__global__ void __launch_bounds__(256) dxfield_kernel(int nx, float *dx, const float *hy,
                                                      int src_pos, float src_val, int src_on)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 1 || i >= nx) return;
    float d = dx[i] + 0.5f * (hy[i - 1] - hy[i]);
    if (src_on && i == src_pos) d = d + src_val;
    dx[i] = d;
}

// ex = nax (dx - ix [- ncx sx]); ix += nbx ex; [sx = ncx sx + ndx ex]
template <bool DEBYE>
__global__ void __launch_bounds__(256) exfield_kernel(int nx, const float *nax, const float *nbx,
                                                      const float *ncx, const float *ndx,
                                                      const float *dx, float *ix, float *sx,
                                                      float *ex)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 1 || i >= nx) return;
    float e;
    if (DEBYE) {
        const float s = sx[i];
        e = nax[i] * (dx[i] - ix[i] - ncx[i] * s);
        ix[i] = ix[i] + nbx[i] * e;
        sx[i] = ncx[i] * s + ndx[i] * e;
    } else {
        e = nax[i] * (dx[i] - ix[i]);
        ix[i] = ix[i] + nbx[i] * e;
    }
    ex[i] = e;
}

struct Twiddle1 {
    int nf;
    float c[FDTD_MAX_FREQ], s[FDTD_MAX_FREQ];
};

__global__ void __launch_bounds__(256) fourier1d_kernel(int nx, Twiddle1 tw, const float *ex,
                                                        int probe, float *r_pt, float *i_pt,
                                                        float *r_in, float *i_in)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && r_in) {
        const float e = ex[probe];
        for (int f = 0; f < tw.nf; ++f) {
            r_in[f] = r_in[f] + tw.c[f] * e;
            i_in[f] = i_in[f] - tw.s[f] * e;
        }
    }
    if (i >= nx) return;
    const float e = ex[i];
    for (int f = 0; f < tw.nf; ++f) {
        r_pt[(long)f * nx + i] = r_pt[(long)f * nx + i] + tw.c[f] * e;
        i_pt[(long)f * nx + i] = i_pt[(long)f * nx + i] - tw.s[f] * e;
    }
}

// Coefficients of fd1d/clang/test_2_3.c:75-90 (FP32 arithmetic; ncx_val =
// exp(-dt/tau) in double on the host, narrowed): free space below nx/2, the
// medium from nx/2 on.
__global__ void __launch_bounds__(256) debye_kernel(int nx, float dt, float chi, float tau,
                                                    float epsr, float sigma, float ncx_val,
                                                    float *nax, float *nbx, float *ncx, float *ndx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx) return;
    const float eps0 = 8.854e-12f;
    if (i < nx / 2) {
        nax[i] = 1.0f;
        nbx[i] = 0.0f;
        if (ncx) { ncx[i] = 0.0f; ndx[i] = 0.0f; }
        return;
    }
    const bool debye = ncx != nullptr;
    const float den = debye ? epsr + sigma * dt / eps0 + chi * dt / tau : epsr + sigma * dt / eps0;
    nax[i] = 1 / den;
    nbx[i] = sigma * dt / eps0;
    if (debye) {
        ncx[i] = ncx_val;
        ndx[i] = chi * dt / tau;
    }
}

// Fused run: each workgroup owns DF_W cells and evolves, for K <= DF_MAXK
// steps, the region of its cells widened by K on both sides (one step moves
// information one cell each way: dx[i] takes hy[i-1], hy[i] takes ex[i+1]),
// so its own cells are exact after K steps.  ex and hy sit in LDS (the only
// fields read across cells), the rest in registers; per step two barriers.
// Reads one state set and writes the other (dx ix sx ex hy bc: a block near
// the right edge may simulate the boundary it does not own); the DFT planes
// and the incident-probe DFT are updated in place by their single owner.
// Operands as the per-step kernels above, so the bits are theirs.
template <int NF>
__global__ void __launch_bounds__(DF_NT) dform_tb_kernel(const DformTbParams P)
{
    constexpr int CPT = (DF_W + 2 * DF_MAXK + DF_NT - 1) / DF_NT;
    __shared__ float ex_s[DF_W + 2 * DF_MAXK], hy_s[DF_W + 2 * DF_MAXK];
    const int nx = P.nx, K = P.K, tid = threadIdx.x;
    const int o0 = blockIdx.x * DF_W, o1 = min(o0 + DF_W, nx);
    const int r0 = max(o0 - K, 0), r1 = min(o1 + K, nx), R = r1 - r0;
    const bool debye = P.ncx != nullptr;
    const bool left_abc = r0 == 0, right_abc = r1 == nx;   // the region holds that grid edge
    float dx[CPT], ix[CPT], sx[CPT], na[CPT], nb[CPT], nc[CPT], nd[CPT];
    float fr[CPT][NF ? NF : 1], fi[CPT][NF ? NF : 1];
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;          // bc of the edge cells' threads
    float rin[NF ? NF : 1], iin[NF ? NF : 1];
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int j = tid + c * DF_NT, i = r0 + j;
        dx[c] = ix[c] = sx[c] = na[c] = nb[c] = nc[c] = nd[c] = 0.f;
#pragma unroll
        for (int f = 0; f < (NF ? NF : 1); ++f) { fr[c][f] = 0.f; fi[c][f] = 0.f; }
        if (j < R) {
            dx[c] = P.dx_i[i]; ix[c] = P.ix_i[i];
            na[c] = P.nax[i]; nb[c] = P.nbx[i];
            if (debye) { sx[c] = P.sx_i[i]; nc[c] = P.ncx[i]; nd[c] = P.ndx[i]; }
            ex_s[j] = P.ex_i[i];
            hy_s[j] = P.hy_i[i];
            if (i >= o0 && i < o1) {
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    if (f < P.nf) {
                        fr[c][f] = P.r_pt[(long)f * nx + i];
                        fi[c][f] = P.i_pt[(long)f * nx + i];
                    }
            }
            if (i == 0 && left_abc) { b0 = P.bc_i[0]; b1 = P.bc_i[1]; }
            if (i == nx - 2 && right_abc) { b2 = P.bc_i[2]; b3 = P.bc_i[3]; }
        }
    }
    const bool own_inc = NF > 0 && P.r_in && o0 <= DF_INC_PROBE && DF_INC_PROBE < o1;
#pragma unroll
    for (int f = 0; f < (NF ? NF : 1); ++f) { rin[f] = 0.f; iin[f] = 0.f; }
    if (own_inc && DF_INC_PROBE - r0 >= 0 && (DF_INC_PROBE - r0) % DF_NT == tid) {
#pragma unroll
        for (int f = 0; f < NF; ++f)
            if (f < P.nf) { rin[f] = P.r_in[f]; iin[f] = P.i_in[f]; }
    }
    __syncthreads();
    for (int s = 0; s < K; ++s) {
        // D update, source, E update, DFT (ex of every cell after the E update)
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int j = tid + c * DF_NT, i = r0 + j;
            if (j >= R) continue;
            if (i >= 1) {
                const float hl = hy_s[j > 0 ? j - 1 : 0];          // j = 0: a cut (stale)
                float d = dx[c] + 0.5f * (hl - hy_s[j]);
                if (i == P.src_pos) d = d + P.src_val[s];
                dx[c] = d;
                float e;
                if (debye) {
                    const float sv = sx[c];
                    e = na[c] * (d - ix[c] - nc[c] * sv);
                    ix[c] = ix[c] + nb[c] * e;
                    sx[c] = nc[c] * sv + nd[c] * e;
                } else {
                    e = na[c] * (d - ix[c]);
                    ix[c] = ix[c] + nb[c] * e;
                }
                ex_s[j] = e;
            }
            const float ev = ex_s[j];
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f < P.nf) {
                    fr[c][f] = fr[c][f] + P.tw_c[s][f] * ev;
                    fi[c][f] = fi[c][f] - P.tw_s[s][f] * ev;
                }
            if (own_inc && i == DF_INC_PROBE) {
#pragma unroll
                for (int f = 0; f < NF; ++f)
                    if (f < P.nf) {
                        rin[f] = rin[f] + P.tw_c[s][f] * ev;
                        iin[f] = iin[f] - P.tw_s[s][f] * ev;
                    }
            }
        }
        __syncthreads();
        // two-step boundary at the grid edges, H update, probes
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int j = tid + c * DF_NT, i = r0 + j;
            if (j >= R || i >= nx - 1) continue;
            float e0 = ex_s[j];
            float e1 = ex_s[j + 1 < R ? j + 1 : j];                // the right cut (stale)
            if (i == 0 && left_abc) {
                const float ob0 = b0;
                ex_s[0] = ob0;
                b0 = b1;
                b1 = e1;
                e0 = ob0;
            }
            if (i == nx - 2 && right_abc) {
                const float ob3 = b3;
                ex_s[j + 1] = ob3;
                b3 = b2;
                b2 = e0;
                e1 = ob3;
            }
            hy_s[j] = hy_s[j] + 0.5f * (e0 - e1);
            if (i >= o0 && i < o1) {
                for (int p = 0; p < P.nprobe; ++p)
                    if (P.probe_k[p] == i)
                        P.probe_out[(long)p * P.probe_len + (P.t_first - 1 + s)] = e0;
            }
            if (i == nx - 2 && right_abc && nx - 1 >= o0 && nx - 1 < o1) {
                for (int p = 0; p < P.nprobe; ++p)
                    if (P.probe_k[p] == nx - 1)
                        P.probe_out[(long)p * P.probe_len + (P.t_first - 1 + s)] = e1;
            }
        }
        __syncthreads();
    }
    // the owned cells go out
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const int j = tid + c * DF_NT, i = r0 + j;
        if (j >= R || i < o0 || i >= o1) continue;
        P.dx_o[i] = dx[c];
        P.ix_o[i] = ix[c];
        if (debye) P.sx_o[i] = sx[c];
        P.ex_o[i] = ex_s[j];
        P.hy_o[i] = hy_s[j];
#pragma unroll
        for (int f = 0; f < NF; ++f)
            if (f < P.nf) {
                P.r_pt[(long)f * nx + i] = fr[c][f];
                P.i_pt[(long)f * nx + i] = fi[c][f];
            }
        if (i == 0) { P.bc_o[0] = b0; P.bc_o[1] = b1; }
        if (i == nx - 2) { P.bc_o[2] = b2; P.bc_o[3] = b3; }
        if (own_inc && i == DF_INC_PROBE) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
                if (f < P.nf) { P.r_in[f] = rin[f]; P.i_in[f] = iin[f]; }
        }
    }
}

inline int nblk(long n) { return (int)((n + 255) / 256); }

}  // namespace

int preload_dform1d() { return preload_kernel(reinterpret_cast<const void *>(&dxfield_kernel)); }

int launch_dxfield(int nx, float *dx, const float *hy, int src_pos, float src_val, int src_on,
                   hipStream_t s)
{
    hipLaunchKernelGGL(dxfield_kernel, dim3(nblk(nx)), dim3(256), 0, s, nx, dx, hy, src_pos,
                       src_val, src_on);
    return hip_check(hipGetLastError(), "dxfield");
}

int launch_exfield_dform(int nx, const fdtd_medium1d &md, const float *dx, float *ix, float *sx,
                         float *ex, hipStream_t s)
{
    if (md.ncx)
        hipLaunchKernelGGL(exfield_kernel<true>, dim3(nblk(nx)), dim3(256), 0, s, nx, md.nax,
                           md.nbx, md.ncx, md.ndx, dx, ix, sx, ex);
    else
        hipLaunchKernelGGL(exfield_kernel<false>, dim3(nblk(nx)), dim3(256), 0, s, nx, md.nax,
                           md.nbx, md.ncx, md.ndx, dx, ix, sx, ex);
    return hip_check(hipGetLastError(), "exfield (D form)");
}

int launch_fourier1d(int t, int nf, int nx, float dt, const float *freq, const float *ex, int probe,
                     const fdtd_ftrans &ft, hipStream_t s)
{
    Twiddle1 tw{};
    tw.nf = nf;
    for (int f = 0; f < nf; ++f) dft_twiddle(t, freq[f], dt, &tw.c[f], &tw.s[f]);
    hipLaunchKernelGGL(fourier1d_kernel, dim3(nblk(nx)), dim3(256), 0, s, nx, tw, ex, probe,
                       ft.r_pt, ft.i_pt, ft.r_in, ft.i_in);
    return hip_check(hipGetLastError(), "fourier (1-D)");
}

int launch_debye(int nx, float dt, float chi, float tau, float epsr, float sigma,
                 const fdtd_medium1d &md, hipStream_t s)
{
    const float ncx_val = (float)exp((double)(-dt / tau));
    hipLaunchKernelGGL(debye_kernel, dim3(nblk(nx)), dim3(256), 0, s, nx, dt, chi, tau, epsr, sigma,
                       ncx_val, md.nax, md.nbx, md.ncx, md.ndx);
    return hip_check(hipGetLastError(), "dielectric (Debye)");
}

int launch_dform_tb(const DformTbParams &p, hipStream_t s)
{
    const dim3 grid((p.nx + DF_W - 1) / DF_W);
    if (p.nf) hipLaunchKernelGGL(dform_tb_kernel<DF_NF>, grid, dim3(DF_NT), 0, s, p);
    else hipLaunchKernelGGL(dform_tb_kernel<0>, grid, dim3(DF_NT), 0, s, p);
    return hip_check(hipGetLastError(), "fd1d_dform_run launch");
}

}  // namespace fdtd
