// Plane-wave total-field/scattered-field (TF/SF) injection and the running
// DFT of the reference's 2-D TM programs with an incident plane wave
// (fd2d/pycuda/fd2d_3_3.py:80-169, fd2d/pycuda/fd2d_3_4.py:121-226; serial
// semantics as fd2d/clang/test_3_4.c:38-147).
//
// The incident wave is a 1-D grid along y (ezi, hxi, ny cells) with the
// two-step absorbing boundary at both ends; the TF/SF box is rows and
// columns npml-1 .. n-npml.  Per step the reference runs ezinct, dfield (which
// also sets the incident source ezi[3]), inctdz, efield, fourier, hxinct,
// hfield, incthx, incthy; the kernels below are those steps, FP32, in the
// reference's operand order without contraction (bit-exact with its C
// program).  The DFT's twiddles cos/sin(2 pi f dt t) are evaluated on the
// host with the host libm, as the C program evaluates them, and passed to
// the kernel by value.
#include <hip/hip_runtime.h>
#include <math.h>

#include "fdtd_internal.h"

namespace fdtd {
namespace {

// ezi[j] += 0.5 (hxi[j-1] - hxi[j]) for j in [1, ny), then the two-step
// boundary at both ends (after every update: one workgroup), then the
// incident source (set_inc).
__global__ void __launch_bounds__(1024) ezinct_kernel(int ny, float *ezi, const float *hxi,
                                                      float *bc, int set_inc, int inc_pos,
                                                      float inc_val)
{
    for (int j = threadIdx.x + 1; j < ny; j += blockDim.x) ezi[j] = ezi[j] + 0.5f * (hxi[j - 1] - hxi[j]);
    __syncthreads();
    if (threadIdx.x == 0) {
        ezi[0] = bc[0];
        bc[0] = bc[1];
        bc[1] = ezi[1];
        ezi[ny - 1] = bc[3];
        bc[3] = bc[2];
        bc[2] = ezi[ny - 2];
        if (set_inc) ezi[inc_pos] = inc_val;
    }
}

__global__ void set_kernel(float *p, int pos, float v) { p[pos] = v; }

// hxi[j] += 0.5 (ezi[j] - ezi[j+1]) for j in [0, ny-1)
__global__ void __launch_bounds__(256) hxinct_kernel(int ny, const float *ezi, float *hxi)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ny - 1) hxi[j] = hxi[j] + 0.5f * (ezi[j] - ezi[j + 1]);
}

// dz at columns npml-1 and ny-npml of rows npml-1 .. nx-npml
__global__ void __launch_bounds__(256) inctdz_kernel(int nx, int ny, int npml, const float *hxi,
                                                     float *dz)
{
    const int i = npml - 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > nx - npml) return;
    float *row = dz + (size_t)i * ny;
    row[npml - 1] = row[npml - 1] + 0.5f * hxi[npml - 2];
    row[ny - npml] = row[ny - npml] - 0.5f * hxi[ny - npml];
}

// hx at columns npml-2 and ny-npml of rows npml-1 .. nx-npml
__global__ void __launch_bounds__(256) incthx_kernel(int nx, int ny, int npml, const float *ezi,
                                                     float *hx)
{
    const int i = npml - 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > nx - npml) return;
    float *row = hx + (size_t)i * ny;
    row[npml - 2] = row[npml - 2] + 0.5f * ezi[npml - 1];
    row[ny - npml] = row[ny - npml] - 0.5f * ezi[ny - npml];
}

// hy at rows npml-2 and nx-npml of columns npml-1 .. ny-npml
__global__ void __launch_bounds__(256) incthy_kernel(int nx, int ny, int npml, const float *ezi,
                                                     float *hy)
{
    const int j = npml - 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j > ny - npml) return;
    float *a = hy + (size_t)(npml - 2) * ny + j;
    float *b = hy + (size_t)(nx - npml) * ny + j;
    *a = *a - 0.5f * ezi[j];
    *b = *b + 0.5f * ezi[j];
}

struct Twiddle {
    int nf;
    float c[FDTD_MAX_FREQ], s[FDTD_MAX_FREQ];
};

// r_pt[f][m] += c_f ez[m], i_pt[f][m] -= s_f ez[m] (frequency-major planes
// of n cells); the first thread also transforms the incident probe
// ezi[probe] into r_in, i_in.
__global__ void __launch_bounds__(256) fourier_kernel(long n, Twiddle tw, const float *ez,
                                                      const float *ezi, int probe, float *r_pt,
                                                      float *i_pt, float *r_in, float *i_in)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0 && r_in) {
        const float e = ezi[probe];
        for (int f = 0; f < tw.nf; ++f) {
            r_in[f] = r_in[f] + tw.c[f] * e;
            i_in[f] = i_in[f] - tw.s[f] * e;
        }
    }
    if (m >= n) return;
    const float e = ez[m];
    for (int f = 0; f < tw.nf; ++f) {
        r_pt[f * n + m] = r_pt[f * n + m] + tw.c[f] * e;
        i_pt[f * n + m] = i_pt[f * n + m] - tw.s[f] * e;
    }
}

inline int blocks(long n, int b) { return (int)((n + b - 1) / b); }

// The incident grid for a chunk of steps (one workgroup: every phase waits
// for the previous), recording per step what the fused 2-D launches read:
// table[s] = {hxi[npml-2], hxi[ny-npml] (before this step's hxinct),
// ezi[npml-1 .. ny-npml] (after ezinct and the source)}; also the incident
// probe's DFT (r_in, i_in; the reference's fourier after efield).
__global__ void __launch_bounds__(1024) inc_chunk_kernel(int ny, float *ezi, float *hxi, float *bc,
                                                         float *r_in, float *i_in, float *table,
                                                         IncChunk a)
{
    const int tid = threadIdx.x, m = a.npml, nbox = ny - 2 * m + 2;
    for (int s = 0; s < a.n; ++s) {
        for (int j = tid + 1; j < ny; j += blockDim.x) ezi[j] = ezi[j] + 0.5f * (hxi[j - 1] - hxi[j]);
        __syncthreads();
        if (tid == 0) {
            ezi[0] = bc[0];
            bc[0] = bc[1];
            bc[1] = ezi[1];
            ezi[ny - 1] = bc[3];
            bc[3] = bc[2];
            bc[2] = ezi[ny - 2];
            ezi[a.inc_pos] = a.wave[s];
        }
        __syncthreads();
        float *row = table + (size_t)s * a.stride;
        if (tid == 0) {
            row[0] = hxi[m - 2];
            row[1] = hxi[ny - m];
            if (r_in) {
                const float e = ezi[a.probe];
                for (int f = 0; f < a.nf; ++f) {
                    r_in[f] = r_in[f] + a.c[s][f] * e;
                    i_in[f] = i_in[f] - a.s[s][f] * e;
                }
            }
        }
        for (int k = tid; k < nbox; k += blockDim.x) row[2 + k] = ezi[m - 1 + k];
        __syncthreads();
        for (int j = tid; j < ny - 1; j += blockDim.x) hxi[j] = hxi[j] + 0.5f * (ezi[j] - ezi[j + 1]);
        __syncthreads();
    }
}

// The same with the incident grid held in LDS for the whole chunk (ezi,
// hxi: 2 ny floats; bc in registers of thread 0): per step only barriers, no
// global round trips; the table rows go out as plain stores.
__global__ void __launch_bounds__(1024) inc_chunk_lds_kernel(int ny, float *ezi, float *hxi,
                                                             float *bc, float *r_in, float *i_in,
                                                             float *table, IncChunk a)
{
    extern __shared__ float sm[];
    float *e = sm, *h = sm + ny;
    const int tid = threadIdx.x, m = a.npml, nbox = ny - 2 * m + 2;
    for (int j = tid; j < ny; j += blockDim.x) { e[j] = ezi[j]; h[j] = hxi[j]; }
    float b0 = 0.f, b1 = 0.f, b2 = 0.f, b3 = 0.f;
    if (tid == 0) { b0 = bc[0]; b1 = bc[1]; b2 = bc[2]; b3 = bc[3]; }
    __syncthreads();
    for (int s = 0; s < a.n; ++s) {
        for (int j = tid + 1; j < ny; j += blockDim.x) e[j] = e[j] + 0.5f * (h[j - 1] - h[j]);
        __syncthreads();
        if (tid == 0) {
            e[0] = b0;
            b0 = b1;
            b1 = e[1];
            e[ny - 1] = b3;
            b3 = b2;
            b2 = e[ny - 2];
            e[a.inc_pos] = a.wave[s];
        }
        __syncthreads();
        float *row = table + (size_t)s * a.stride;
        if (tid == 0) {
            row[0] = h[m - 2];
            row[1] = h[ny - m];
            if (r_in) {
                const float v = e[a.probe];
                for (int f = 0; f < a.nf; ++f) {
                    r_in[f] = r_in[f] + a.c[s][f] * v;
                    i_in[f] = i_in[f] - a.s[s][f] * v;
                }
            }
        }
        for (int k = tid; k < nbox; k += blockDim.x) row[2 + k] = e[m - 1 + k];
        __syncthreads();
        for (int j = tid; j < ny - 1; j += blockDim.x) h[j] = h[j] + 0.5f * (e[j] - e[j + 1]);
        __syncthreads();
    }
    for (int j = tid; j < ny; j += blockDim.x) { ezi[j] = e[j]; hxi[j] = h[j]; }
    if (tid == 0) { bc[0] = b0; bc[1] = b1; bc[2] = b2; bc[3] = b3; }
}

}  // namespace

int preload_tfsf() { return preload_kernel(reinterpret_cast<const void *>(&set_kernel)); }

int launch_inc_chunk(int ny, float *ezi, float *hxi, float *bc, float *r_in, float *i_in,
                     float *table, const IncChunk &a, hipStream_t s)
{
    const size_t lds = 2 * (size_t)ny * sizeof(float);
    if (lds <= 64 * 1024)
        hipLaunchKernelGGL(inc_chunk_lds_kernel, dim3(1), dim3(1024), lds, s, ny, ezi, hxi, bc, r_in,
                           i_in, table, a);
    else
        hipLaunchKernelGGL(inc_chunk_kernel, dim3(1), dim3(1024), 0, s, ny, ezi, hxi, bc, r_in, i_in,
                           table, a);
    return hip_check(hipGetLastError(), "incident chunk");
}

// The reference's twiddle: cosf/sinf of 2*M_PI*freq*dt*t evaluated in double
// and narrowed to float (fd2d/clang/test_3_4.c:42-43).
void dft_twiddle(int t, float freq, float dt, float *c, float *s)
{
    const float a = (float)(2 * M_PI * (double)freq * (double)dt * (double)t);
    *c = cosf(a);
    *s = sinf(a);
}

int launch_ezinct(int ny, float *ezi, const float *hxi, float *bc, int set_inc, int inc_pos,
                  float inc_val, hipStream_t s)
{
    hipLaunchKernelGGL(ezinct_kernel, dim3(1), dim3(1024), 0, s, ny, ezi, hxi, bc, set_inc, inc_pos,
                       inc_val);
    return hip_check(hipGetLastError(), "ezinct");
}

int launch_set(float *p, int pos, float v, hipStream_t s)
{
    hipLaunchKernelGGL(set_kernel, dim3(1), dim3(1), 0, s, p, pos, v);
    return hip_check(hipGetLastError(), "incident source");
}

int launch_hxinct(int ny, const float *ezi, float *hxi, hipStream_t s)
{
    hipLaunchKernelGGL(hxinct_kernel, dim3(blocks(ny, 256)), dim3(256), 0, s, ny, ezi, hxi);
    return hip_check(hipGetLastError(), "hxinct");
}

int launch_inctdz(int nx, int ny, int npml, const float *hxi, float *dz, hipStream_t s)
{
    hipLaunchKernelGGL(inctdz_kernel, dim3(blocks(nx - 2 * npml + 2, 256)), dim3(256), 0, s, nx, ny,
                       npml, hxi, dz);
    return hip_check(hipGetLastError(), "inctdz");
}

int launch_incthx(int nx, int ny, int npml, const float *ezi, float *hx, hipStream_t s)
{
    hipLaunchKernelGGL(incthx_kernel, dim3(blocks(nx - 2 * npml + 2, 256)), dim3(256), 0, s, nx, ny,
                       npml, ezi, hx);
    return hip_check(hipGetLastError(), "incthx");
}

int launch_incthy(int nx, int ny, int npml, const float *ezi, float *hy, hipStream_t s)
{
    hipLaunchKernelGGL(incthy_kernel, dim3(blocks(ny - 2 * npml + 2, 256)), dim3(256), 0, s, nx, ny,
                       npml, ezi, hy);
    return hip_check(hipGetLastError(), "incthy");
}

int launch_fourier(int t, int nf, int nx, int ny, float dt, const float *freq, const float *ezi,
                   int probe, const float *ez, const fdtd_ftrans &ft, hipStream_t s)
{
    Twiddle tw{};
    tw.nf = nf;
    for (int f = 0; f < nf; ++f) dft_twiddle(t, freq[f], dt, &tw.c[f], &tw.s[f]);
    const long n = (long)nx * ny;
    hipLaunchKernelGGL(fourier_kernel, dim3(blocks(n, 256)), dim3(256), 0, s, n, tw, ez, ezi, probe,
                       ft.r_pt, ft.i_pt, ft.r_in, ft.i_in);
    return hip_check(hipGetLastError(), "fourier");
}

}  // namespace fdtd
