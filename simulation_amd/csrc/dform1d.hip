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

inline int nblk(long n) { return (int)((n + 255) / 256); }

}  // namespace

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

}  // namespace fdtd
