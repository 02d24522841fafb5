// fd1d.hip -- 1-D Ex/Hy kernels for gfx950.
//
// Semantics: the serial order of fd1d/clang/fd1d_1_2.c:151-167 and
// fd1d/clang/fd1d_1_5.c:25-41 (E update, source, two-step boundary, H update),
// in the expression order of the PyCUDA kernels fd1d/pycuda/fd1d_1_5.py:49-67.
// Built with -ffp-contract=off: every multiply and add rounds separately.
#include "fdtd_internal.h"

namespace fdtd {
namespace {

// E update of one cell.  ca == NULL means ca = 1 and cb == NULL means cb = 0.5:
// ca*ex == ex exactly for ca = 1, so one body covers the += forms of
// fd1d/pycuda/fd1d_1_2.py:47 and fd1d/pycuda/fd1d_1_3.py:50.
template <typename T>
__device__ __forceinline__ T e_cell(T e, T hl, T hr, const T *ca, const T *cb, int i)
{
    T d = hl - hr;
    if (ca) return ca[i] * e + cb[i] * d;
    if (cb) return e + cb[i] * d;
    return e + T(0.5) * d;
}

template <typename T>
__global__ void fd1d_exfield_kernel(int nx, const T *__restrict__ ca, const T *__restrict__ cb,
                                    T *__restrict__ ex, const T *__restrict__ hy,
                                    int src_mode, int src_pos, T src_val)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx; i += blockDim.x * gridDim.x) {
        T v = ex[i];
        bool changed = false;
        if (i >= 1) { v = e_cell(v, hy[i - 1], hy[i], ca, cb, i); changed = true; }
        // the owner of the source cell applies the source after its update
        if (i == src_pos && src_mode == FDTD_SRC_HARD) { v = src_val; changed = true; }
        else if (i == src_pos && src_mode == FDTD_SRC_ADD) { v = v + src_val; changed = true; }
        if (changed) ex[i] = v;
    }
}

// Boundary + H update.  Thread i = 0 owns the low-end boundary (reads bc[0],
// bc[1], ex[1]; writes ex[0], bc[0], bc[1]) and thread i = nx-2 the high end
// (reads bc[2], bc[3], ex[nx-2]; writes ex[nx-1], bc[2], bc[3]); no other
// thread touches those words, and the H update of those two threads uses the
// new boundary values, as the serial program does.
template <typename T>
__global__ void fd1d_hyfield_kernel(int nx, T *__restrict__ ex, T *__restrict__ hy,
                                    T *__restrict__ bc)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx - 1; i += blockDim.x * gridDim.x) {
        T e0 = ex[i], e1 = ex[i + 1];
        if (bc) {
            if (i == 0) {
                T b0 = bc[0], b1 = bc[1];
                ex[0] = b0; bc[0] = b1; bc[1] = e1;
                e0 = b0;
            }
            if (i == nx - 2) {
                T b3 = bc[3], b2 = bc[2];
                ex[nx - 1] = b3; bc[3] = b2; bc[2] = e0;
                e1 = b3;
            }
        }
        hy[i] = hy[i] + T(0.5) * (e0 - e1);
    }
}

// All steps in one workgroup: ex and hy live in LDS, three barriers a step.
struct ProbeK { int n; int k[FDTD_MAX_PROBES]; };

template <typename T>
__global__ void __launch_bounds__(1024)
fd1d_run_kernel(int nx, int nsteps, const T *__restrict__ ca, const T *__restrict__ cb,
                T *__restrict__ ex_g, T *__restrict__ hy_g, T *__restrict__ bc_g,
                int src_mode, int src_pos, const T *__restrict__ src_wave,
                ProbeK pk, T *__restrict__ probe_out)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *ex = reinterpret_cast<T *>(smem);
    T *hy = ex + nx;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < nx; i += nt) { ex[i] = ex_g[i]; hy[i] = hy_g[i]; }
    T b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    if (tid == 0 && bc_g) { b0 = bc_g[0]; b1 = bc_g[1]; b2 = bc_g[2]; b3 = bc_g[3]; }
    __syncthreads();
    for (int s = 0; s < nsteps; ++s) {
        const T sv = src_wave ? src_wave[s] : T(0);
        for (int i = tid; i < nx; i += nt) {
            T v = ex[i];
            if (i >= 1) v = e_cell(v, hy[i - 1], hy[i], ca, cb, i);
            if (i == src_pos && src_mode == FDTD_SRC_HARD) v = sv;
            else if (i == src_pos && src_mode == FDTD_SRC_ADD) v = v + sv;
            ex[i] = v;
        }
        __syncthreads();
        if (tid == 0) {
            if (bc_g) {
                ex[0] = b0; b0 = b1; b1 = ex[1];
                ex[nx - 1] = b3; b3 = b2; b2 = ex[nx - 2];
            }
            for (int p = 0; p < pk.n; ++p) probe_out[(size_t)p * nsteps + s] = ex[pk.k[p]];
        }
        __syncthreads();
        for (int i = tid; i < nx - 1; i += nt) hy[i] = hy[i] + T(0.5) * (ex[i] - ex[i + 1]);
        __syncthreads();
    }
    for (int i = tid; i < nx; i += nt) { ex_g[i] = ex[i]; hy_g[i] = hy[i]; }
    if (tid == 0 && bc_g) { bc_g[0] = b0; bc_g[1] = b1; bc_g[2] = b2; bc_g[3] = b3; }
}

// ca, cb of fd1d/pycuda/fd1d_1_5.py:71-78, in FP64, stored as T.
template <typename T>
__global__ void fd1d_dielectric_kernel(int nx, double dt, double epsr, double sigma,
                                       T *__restrict__ ca, T *__restrict__ cb)
{
    const double eps0 = 8.854e-12;
    const double epsf = dt * sigma / (2 * eps0 * epsr);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx; i += blockDim.x * gridDim.x) {
        double a = 1.0, b = 0.5;
        if (i >= nx / 2) { a = (1 - epsf) / (1 + epsf); b = 0.5 / (epsr * (1 + epsf)); }
        ca[i] = (T)a;
        cb[i] = (T)b;
    }
}

constexpr int kLdsLimit = 64 * 1024;

inline int grid1d(int n, int block) { int g = (n + block - 1) / block; return g < 1 ? 1 : (g > 65535 ? 65535 : g); }

}  // namespace

int fd1d_run_lds_limit() { return kLdsLimit; }

int preload_fd1d() { return preload_kernel(reinterpret_cast<const void *>(&fd1d_dielectric_kernel<float>)); }

template <typename T>
int launch_fd1d_exfield(int nx, const T *ca, const T *cb, T *ex, const T *hy,
                        int src_mode, int src_pos, T src_val, hipStream_t s)
{
    hipLaunchKernelGGL(fd1d_exfield_kernel<T>, dim3(grid1d(nx, 256)), dim3(256), 0, s,
                       nx, ca, cb, ex, hy, src_mode, src_pos, src_val);
    return hip_check(hipGetLastError(), "fd1d_exfield launch");
}

template <typename T>
int launch_fd1d_hyfield(int nx, T *ex, T *hy, T *bc, hipStream_t s)
{
    hipLaunchKernelGGL(fd1d_hyfield_kernel<T>, dim3(grid1d(nx - 1, 256)), dim3(256), 0, s,
                       nx, ex, hy, bc);
    return hip_check(hipGetLastError(), "fd1d_hyfield launch");
}

template <typename T>
int launch_fd1d_run(int nx, int nsteps, const T *ca, const T *cb, T *ex, T *hy, T *bc,
                    int src_mode, int src_pos, const T *src_wave,
                    int nprobe, const int *probe_k, T *probe_out, hipStream_t s)
{
    ProbeK pk{};
    pk.n = nprobe;
    for (int p = 0; p < nprobe; ++p) pk.k[p] = probe_k[p];
    int nt = ((nx + 63) / 64) * 64;
    if (nt > 1024) nt = 1024;
    const size_t lds = 2 * sizeof(T) * (size_t)nx;
    hipLaunchKernelGGL(fd1d_run_kernel<T>, dim3(1), dim3(nt), lds, s,
                       nx, nsteps, ca, cb, ex, hy, bc, src_mode, src_pos, src_wave, pk, probe_out);
    return hip_check(hipGetLastError(), "fd1d_run launch");
}

template <typename T>
int launch_fd1d_dielectric(int nx, double dt, double epsr, double sigma, T *ca, T *cb,
                           hipStream_t s)
{
    hipLaunchKernelGGL(fd1d_dielectric_kernel<T>, dim3(grid1d(nx, 256)), dim3(256), 0, s,
                       nx, dt, epsr, sigma, ca, cb);
    return hip_check(hipGetLastError(), "fd1d_dielectric launch");
}

template int launch_fd1d_dielectric<float>(int, double, double, double, float *, float *, hipStream_t);
template int launch_fd1d_dielectric<double>(int, double, double, double, double *, double *, hipStream_t);

template int launch_fd1d_exfield<float>(int, const float *, const float *, float *, const float *, int, int, float, hipStream_t);
template int launch_fd1d_exfield<double>(int, const double *, const double *, double *, const double *, int, int, double, hipStream_t);
template int launch_fd1d_hyfield<float>(int, float *, float *, float *, hipStream_t);
template int launch_fd1d_hyfield<double>(int, double *, double *, double *, hipStream_t);
template int launch_fd1d_run<float>(int, int, const float *, const float *, float *, float *, float *, int, int, const float *, int, const int *, float *, hipStream_t);
template int launch_fd1d_run<double>(int, int, const double *, const double *, double *, double *, double *, int, int, const double *, int, const int *, double *, hipStream_t);

}  // namespace fdtd
