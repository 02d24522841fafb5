// fd2d.hip -- 2-D TM kernels for gfx950: the per-step D/E/H updates with the
// reference's argument order, the FP64 coefficient set-up, probes and the
// FP64 energy reduction.
//
// Update arithmetic follows the PyCUDA kernels operand for operand
// (fd2d/pycuda/fd2d_3_2.py:219-260, fd2d/pycuda/fd2d_3_4.py:175-185); the
// library is built with -ffp-contract=off so each product and sum rounds on
// its own, as in the reference's C programs.
#include <mutex>

#include "fdtd_internal.h"

namespace fdtd {
namespace {

// 64 lanes along j (one wavefront reads 256 contiguous bytes of a row),
// 4 rows per block.
constexpr int BX = 64, BY = 4;

inline dim3 grid2d(int nx, int ny)
{
    int gx = (ny + BX - 1) / BX;
    int gy = (nx + BY - 1) / BY;
    if (gy > 65535) gy = 65535;       // kernels stride over i beyond this
    return dim3(gx < 1 ? 1 : gx, gy < 1 ? 1 : gy);
}

template <bool PML>
__global__ void __launch_bounds__(BX * BY)
fd2d_dfield_kernel(int nx, int ny, PmlPtrs pml, float *__restrict__ dz,
                   const float *__restrict__ hx, const float *__restrict__ hy, SrcList src)
{
    const int j = blockIdx.x * BX + threadIdx.x;
    if (j >= ny) return;
    for (int i = blockIdx.y * BY + threadIdx.y; i < nx; i += BY * gridDim.y) {
        const size_t n = (size_t)i * ny + j;
        bool changed = false;
        float v = 0.f;
        if (i >= 1 && j >= 1) {
            v = dz[n];
            const float curl = hy[n] - hy[n - ny] - hx[n] + hx[n - 1];
            if (PML) v = pml.gx3[i] * pml.gy3[j] * v + pml.gx2[i] * pml.gy2[j] * 0.5f * curl;
            else v = v + 0.5f * curl;
            changed = true;
        }
        for (int k = 0; k < src.n; ++k)
            if (i == src.i[k] && j == src.j[k]) {
                v = src.wave ? src.wave[(size_t)k * src.wave_len + (src.t - 1)] : src.v[k];
                changed = true;
            }
        if (changed) dz[n] = v;
    }
}

template <bool LOSSY>
__global__ void __launch_bounds__(256)
fd2d_efield_kernel(size_t ncell, const float *__restrict__ naz, const float *__restrict__ nbz,
                   const float *__restrict__ dz, float *__restrict__ iz, float *__restrict__ ez)
{
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < ncell;
         n += (size_t)blockDim.x * gridDim.x) {
        if (LOSSY) {
            const float i0 = iz[n];
            const float e = naz[n] * (dz[n] - i0);
            ez[n] = e;
            iz[n] = i0 + nbz[n] * e;
        } else {
            ez[n] = naz[n] * dz[n];
        }
    }
}

template <bool PML>
__global__ void __launch_bounds__(BX * BY)
fd2d_hfield_kernel(int nx, int ny, PmlPtrs pml, const float *__restrict__ ez,
                   float *__restrict__ ihx, float *__restrict__ ihy,
                   float *__restrict__ hx, float *__restrict__ hy)
{
    const int j = blockIdx.x * BX + threadIdx.x;
    if (j >= ny - 1) return;
    for (int i = blockIdx.y * BY + threadIdx.y; i < nx - 1; i += BY * gridDim.y) {
        const size_t n = (size_t)i * ny + j;
        const float e = ez[n], eR = ez[n + 1], eD = ez[n + ny];
        if (PML) {
            const float ax = ihx[n] + (e - eR);
            const float ay = ihy[n] + (e - eD);
            ihx[n] = ax;
            ihy[n] = ay;
            hx[n] = pml.fy3[j] * hx[n] + pml.fy2[j] * (0.5f * e - 0.5f * eR + pml.fx1[i] * ax);
            hy[n] = pml.fx3[i] * hy[n] - pml.fx2[i] * (0.5f * e - 0.5f * eD + pml.fy1[j] * ay);
        } else {
            hx[n] = hx[n] + 0.5f * (e - eR);
            hy[n] = hy[n] - 0.5f * (e - eD);
        }
    }
}

struct ProbeList { int n; int i[FDTD_MAX_PROBES]; int j[FDTD_MAX_PROBES]; };

__global__ void fd2d_probe_kernel(int ny, const float *__restrict__ ez, ProbeList pl,
                                  float *__restrict__ out, int stride, int col)
{
    const int p = threadIdx.x;
    if (p < pl.n) out[(size_t)p * stride + col] = ez[(size_t)pl.i[p] * ny + pl.j[p]];
}

// x^3 rounded once: x*x = p + pe and p*x = c + ce exactly (fused multiply-add
// error terms), so c + (ce + pe*x) is the correctly rounded cube except within
// ~2^-100 of a rounding tie.  The reference forms the cube with libm pow on the
// host (fd2d/pycuda/fd2d_3_2.py:267-268).
__device__ __forceinline__ double cube_rn(double x)
{
    const double p = x * x;
    const double pe = fma(x, x, -p);
    const double c = p * x;
    const double ce = fma(p, x, -c);
    return c + (ce + pe * x);
}

// One profile set of length len (the x set with len = nx, the y set with
// len = ny).  The reference assigns, for n = 0..npml-1 in order, the f-profile
// of n at indices n and len-2-n and the g-profile at n and len-1-n; the last
// assignment wins, i.e. an index takes the largest such n.
__device__ void pml_index(int idx, int len, int npml, float *f1, float *f2, float *f3,
                          float *g2, float *g3)
{
    int nf = (idx < npml) ? idx : -1;
    const int nf2 = len - 2 - idx;
    if (nf2 >= 0 && nf2 < npml && nf2 > nf) nf = nf2;
    int ng = (idx < npml) ? idx : -1;
    const int ng2 = len - 1 - idx;
    if (ng2 >= 0 && ng2 < npml && ng2 > ng) ng = ng2;
    double a1 = 0.0, a2 = 1.0, a3 = 1.0, b2 = 1.0, b3 = 1.0;
    if (nf >= 0) {
        const double xn = 0.33 * cube_rn((npml - nf - 0.5) / npml);
        a1 = xn;
        a2 = 1 / (1 + xn);
        a3 = (1 - xn) / (1 + xn);
    }
    if (ng >= 0) {
        const double xm = 0.33 * cube_rn((double)(npml - ng) / npml);
        b2 = 1 / (1 + xm);
        b3 = (1 - xm) / (1 + xm);
    }
    f1[idx] = (float)a1; f2[idx] = (float)a2; f3[idx] = (float)a3;
    g2[idx] = (float)b2; g3[idx] = (float)b3;
}

__global__ void fd2d_pmlparam_kernel(int nx, int ny, int npml, fdtd_pml pml)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nx) pml_index(idx, nx, npml, pml.fx1, pml.fx2, pml.fx3, pml.gx2, pml.gx3);
    if (idx < ny) pml_index(idx, ny, npml, pml.fy1, pml.fy2, pml.fy3, pml.gy2, pml.gy3);
}

// The Numba dielectric of fd2d/pycuda/fd2d_3_4.py:233-254 in FP64, one cell
// per thread, generalised to a list of cylinders (first containing cylinder
// wins for each sub-sample); int_sub selects the integer sub-sampling of
// fd2d/clang/fd2d_3_4.c:160-163.
__global__ void __launch_bounds__(256)
fd2d_dielectric_kernel(int nx, int ny, int npml, double dt, CylList cyl, int int_sub,
                       float *__restrict__ naz, float *__restrict__ nbz)
{
    const double eps0 = 8.854e-12;
    const size_t ncell = (size_t)nx * ny;
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < ncell;
         n += (size_t)blockDim.x * gridDim.x) {
        const int i = (int)(n / ny), j = (int)(n % ny);
        double a = 1.0, b = 0.0;
        if (i >= npml && i < nx - npml && j >= npml && j < ny - npml) {
            double epsn = 1.0, cond = 0.0;
            for (int m = -1; m <= 1; ++m) {
                for (int q = -1; q <= 1; ++q) {
                    for (int c = 0; c < cyl.n; ++c) {
                        double x, y;
                        if (int_sub) {
                            x = (double)(nx / 2 - 1 - i + m / 3);
                            y = (double)(ny / 2 - 1 - j + q / 3);
                        } else {
                            // bounding-box rejection; cannot change a result
                            if (fabs(cyl.cx[c] - i) > cyl.rad[c] + 1.0 ||
                                fabs(cyl.cy[c] - j) > cyl.rad[c] + 1.0) continue;
                            x = cyl.cx[c] - i + m / 3.0;
                            y = cyl.cy[c] - j + q / 3.0;
                        }
                        const double d = sqrt(x * x + y * y);
                        if (d <= cyl.rad[c]) {
                            epsn += (cyl.epsr[c] - 1) / 9;
                            cond += cyl.sigma[c] / 9;
                            break;
                        }
                    }
                }
            }
            a = 1 / (epsn + cond * dt / eps0);
            b = cond * dt / eps0;
        }
        naz[n] = (float)a;
        nbz[n] = (float)b;
    }
}

// Deterministic FP64 energy: a fixed grid of blocks each sums a fixed set of
// cells, wave64 shuffles then LDS reduce the block, a second single-block pass
// sums the partials in index order.
constexpr int EB = 1024;   // partial sums (blocks of the first pass)
constexpr int ET = 256;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ double block_sum(double v)
{
    __shared__ double part[ET / 64];
    v = wave_sum(v);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) part[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < ET / 64; ++k) r += part[k];
    return r;
}

__global__ void __launch_bounds__(ET)
fd2d_energy_partial(int ny, int i0, int i1, int j0, int j1, const float *__restrict__ ez,
                    const float *__restrict__ hx, const float *__restrict__ hy,
                    double *__restrict__ work)
{
    const int wj = j1 - j0;
    const size_t ncell = (size_t)(i1 - i0) * wj;
    double acc = 0.0;
    for (size_t k = (size_t)blockIdx.x * ET + threadIdx.x; k < ncell; k += (size_t)EB * ET) {
        const size_t n = (size_t)(i0 + k / wj) * ny + (j0 + k % wj);
        const double e = ez[n], a = hx[n], b = hy[n];
        acc += e * e + a * a + b * b;
    }
    const double r = block_sum(acc);
    if (threadIdx.x == 0) work[1 + blockIdx.x] = r;
}

// Energy of the difference of two field sets over equal windows: window
// (ia0, ja0) of grid a (row length nya) against (ib0, jb0) of grid b.
__global__ void __launch_bounds__(ET)
fd2d_energy_diff_partial(int wn, int wm, int nya, int ia0, int ja0, const float *__restrict__ eza,
                         const float *__restrict__ hxa, const float *__restrict__ hya, int nyb,
                         int ib0, int jb0, const float *__restrict__ ezb,
                         const float *__restrict__ hxb, const float *__restrict__ hyb,
                         double *__restrict__ work)
{
    const size_t ncell = (size_t)wn * wm;
    double acc = 0.0;
    for (size_t k = (size_t)blockIdx.x * ET + threadIdx.x; k < ncell; k += (size_t)EB * ET) {
        const size_t r = k / wm, c = k % wm;
        const size_t na = (ia0 + r) * nya + (ja0 + c), nb = (ib0 + r) * nyb + (jb0 + c);
        const double e = (double)eza[na] - (double)ezb[nb];
        const double a = (double)hxa[na] - (double)hxb[nb];
        const double b = (double)hya[na] - (double)hyb[nb];
        acc += e * e + a * a + b * b;
    }
    const double r = block_sum(acc);
    if (threadIdx.x == 0) work[1 + blockIdx.x] = r;
}

// Sums the partials work[1 .. EB] in index order into *out (out = work for
// the result in work[0]).
__global__ void __launch_bounds__(ET) fd2d_energy_final(const double *work, double *out)
{
    double acc = 0.0;
    for (int k = threadIdx.x; k < EB; k += ET) acc += work[1 + k];
    const double r = block_sum(acc);
    if (threadIdx.x == 0) out[0] = r;
}

}  // namespace

int preload_fd2d() { return preload_kernel(reinterpret_cast<const void *>(&fd2d_pmlparam_kernel)); }

int launch_fd2d_dfield(int nx, int ny, const PmlPtrs *pml, float *dz, const float *hx,
                       const float *hy, const SrcList &src, hipStream_t s)
{
    if (pml)
        hipLaunchKernelGGL(fd2d_dfield_kernel<true>, grid2d(nx, ny), dim3(BX, BY), 0, s,
                           nx, ny, *pml, dz, hx, hy, src);
    else
        hipLaunchKernelGGL(fd2d_dfield_kernel<false>, grid2d(nx, ny), dim3(BX, BY), 0, s,
                           nx, ny, PmlPtrs{}, dz, hx, hy, src);
    return hip_check(hipGetLastError(), "fd2d_dfield launch");
}

int launch_fd2d_efield(int nx, int ny, const float *naz, const float *nbz, const float *dz,
                       float *iz, float *ez, hipStream_t s)
{
    const size_t ncell = (size_t)nx * ny;
    size_t g = (ncell + 255) / 256;
    if (g > 1048576) g = 1048576;
    if (nbz)
        hipLaunchKernelGGL(fd2d_efield_kernel<true>, dim3((unsigned)g), dim3(256), 0, s,
                           ncell, naz, nbz, dz, iz, ez);
    else
        hipLaunchKernelGGL(fd2d_efield_kernel<false>, dim3((unsigned)g), dim3(256), 0, s,
                           ncell, naz, nbz, dz, iz, ez);
    return hip_check(hipGetLastError(), "fd2d_efield launch");
}

int launch_fd2d_hfield(int nx, int ny, const PmlPtrs *pml, const float *ez, float *ihx,
                       float *ihy, float *hx, float *hy, hipStream_t s)
{
    if (pml)
        hipLaunchKernelGGL(fd2d_hfield_kernel<true>, grid2d(nx, ny), dim3(BX, BY), 0, s,
                           nx, ny, *pml, ez, ihx, ihy, hx, hy);
    else
        hipLaunchKernelGGL(fd2d_hfield_kernel<false>, grid2d(nx, ny), dim3(BX, BY), 0, s,
                           nx, ny, PmlPtrs{}, ez, ihx, ihy, hx, hy);
    return hip_check(hipGetLastError(), "fd2d_hfield launch");
}

int launch_fd2d_probe(int ny, const float *ez, int nprobe, const int *pi, const int *pj,
                      float *out, int stride, int col, hipStream_t s)
{
    ProbeList pl{};
    pl.n = nprobe;
    for (int p = 0; p < nprobe; ++p) { pl.i[p] = pi[p]; pl.j[p] = pj[p]; }
    hipLaunchKernelGGL(fd2d_probe_kernel, dim3(1), dim3(64), 0, s, ny, ez, pl, out, stride, col);
    return hip_check(hipGetLastError(), "probe launch");
}

int launch_fd2d_pmlparam(int nx, int ny, int npml, const fdtd_pml &pml, hipStream_t s)
{
    const int n = nx > ny ? nx : ny;
    hipLaunchKernelGGL(fd2d_pmlparam_kernel, dim3((n + 255) / 256), dim3(256), 0, s,
                       nx, ny, npml, pml);
    return hip_check(hipGetLastError(), "fd2d_pmlparam launch");
}

int launch_fd2d_dielectric(int nx, int ny, int npml, double dt, const CylList &cyl,
                           int int_sub, float *naz, float *nbz, hipStream_t s)
{
    const size_t ncell = (size_t)nx * ny;
    size_t g = (ncell + 255) / 256;
    if (g > 262144) g = 262144;
    hipLaunchKernelGGL(fd2d_dielectric_kernel, dim3((unsigned)g), dim3(256), 0, s,
                       nx, ny, npml, dt, cyl, int_sub, naz, nbz);
    return hip_check(hipGetLastError(), "fd2d_dielectric launch");
}

int launch_fd2d_energy(int nx, int ny, int i0, int i1, int j0, int j1, const float *ez,
                       const float *hx, const float *hy, double *work, hipStream_t s)
{
    (void)nx;
    hipLaunchKernelGGL(fd2d_energy_partial, dim3(EB), dim3(ET), 0, s,
                       ny, i0, i1, j0, j1, ez, hx, hy, work);
    int rc = hip_check(hipGetLastError(), "fd2d_energy launch");
    if (rc) return rc;
    hipLaunchKernelGGL(fd2d_energy_final, dim3(1), dim3(ET), 0, s, work, work);
    return hip_check(hipGetLastError(), "fd2d_energy launch");
}

int launch_fd2d_energy_diff(int wn, int wm, int nya, int ia0, int ja0, const float *eza,
                            const float *hxa, const float *hya, int nyb, int ib0, int jb0,
                            const float *ezb, const float *hxb, const float *hyb, double *work,
                            hipStream_t s)
{
    hipLaunchKernelGGL(fd2d_energy_diff_partial, dim3(EB), dim3(ET), 0, s, wn, wm, nya, ia0, ja0,
                       eza, hxa, hya, nyb, ib0, jb0, ezb, hxb, hyb, work);
    int rc = hip_check(hipGetLastError(), "fd2d_energy_diff launch");
    if (rc) return rc;
    hipLaunchKernelGGL(fd2d_energy_final, dim3(1), dim3(ET), 0, s, work, work);
    return hip_check(hipGetLastError(), "fd2d_energy_diff launch");
}

// ------------------------------------------------ energy without a workspace
// The partial sums of fd2d_energy_noalloc live in this file-scope device
// array (one copy per GPU, held until the process ends).  Host code never
// passes `scratch` itself to a kernel (in host code that name is a host-side
// stand-in): it passes the device address hipGetSymbolAddress returns for the
// current device, looked up once per device and kept.
static __device__ __align__(256) double scratch[FDTD_ENERGY_WORK];

int energy_scratch(double **out)
{
    static std::mutex mu;
    static double *addr[64] = {};
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return fail("device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
        if (rc) return rc;
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return 0;
}

size_t energy_scratch_elems() { return FDTD_ENERGY_WORK; }

int launch_fd2d_energy_to(int nx, int ny, int i0, int i1, int j0, int j1, const float *ez,
                          const float *hx, const float *hy, double *work, double *out,
                          hipStream_t s)
{
    (void)nx;
    static_assert(1 + EB <= FDTD_ENERGY_WORK, "partials exceed the work area");
    hipLaunchKernelGGL(fd2d_energy_partial, dim3(EB), dim3(ET), 0, s,
                       ny, i0, i1, j0, j1, ez, hx, hy, work);
    int rc = hip_check(hipGetLastError(), "fd2d_energy launch");
    if (rc) return rc;
    hipLaunchKernelGGL(fd2d_energy_final, dim3(1), dim3(ET), 0, s, work, out);
    return hip_check(hipGetLastError(), "fd2d_energy launch");
}

}  // namespace fdtd
