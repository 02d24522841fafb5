/*
 * fdtd_oracle.c -- CPU oracle for the FDTD hot path of dsarvan/simulation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker of the HIP
 * path and the "port" CPU baseline.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it (through oracle/oracle.py); the product
 * (simulation_amd/, libfdtd_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C99 + OpenMP, the serial semantics of:
 *   - the 1-D Ex/Hy programs (fd1d/clang/fd1d_1_2.c, fd1d_1_5.c, test_1_*.c),
 *   - the 2-D TM programs (fd2d/clang/fd2d_3_1.c, test_3_4.c) and the PyCUDA
 *     host set-up they are driven with (pmlparam fd2d/pycuda/fd2d_3_2.py:264-273,
 *     Numba dielectric fd2d/pycuda/fd2d_3_4.py:233-254).
 * Build: oracle/Makefile (-O2 -ffp-contract=off -fopenmp, no -march): with no
 * fused multiply-add the float instance reproduces the reference's FP32 C
 * programs bit for bit, and the double instance its FP64 programs.
 *
 * How each function is pinned is listed in oracle/README.md.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

/* ------------------------------------------------------------ waveforms */

/* fd1d/clang/test_1_2.c:13-15, fd2d/clang/test_3_1.c:13-15 (FP32). */
float orc_gaussian_f32(int t, int t0, float sigma)
{
    return expf(-0.5f*(t - t0)/sigma*(t - t0)/sigma);
}

/* fd1d/clang/fd1d_1_2.c:146-148, fd2d/clang/fd2d_3_1.c:13-15 (FP64). */
double orc_gaussian_f64(int t, int t0, double sigma)
{
    return exp(-0.5*(t - t0)/sigma*(t - t0)/sigma);
}

/* fd2d/clang/test_3_2.c:22-25, fd1d/clang/test_1_5.c (FP32; the argument is
 * formed in double because M_PI is double, then narrowed for sinf). */
float orc_sinusoidal_f32(int t, float ds, float freq)
{
    float dt = ds/6e8f;
    return sinf(2*M_PI*freq*dt*t);
}

/* fd2d/clang/fd2d_3_2.c:22-25, fd1d/clang/fd1d_1_5.c:19-22 (FP64). */
double orc_sinusoidal_f64(int t, double ds, double freq)
{
    double dt = ds/6e8;
    return sin(2*M_PI*freq*dt*t);
}

/* ------------------------------------------------------- 1-D dielectric */

/* ca/cb of a lossy half space i >= nx/2, in double, as
 * fd1d/clang/fd1d_1_5.c:44-55 and fd1d/pycuda/fd1d_1_5.py:71-78
 * (the caller narrows to float where the reference stores FP32). */
void orc_fd1d_dielectric_f64(int nx, double dt, double epsr, double sigma,
                             double *ca, double *cb)
{
    for (int i = 0; i < nx; i++) { ca[i] = 1.0; cb[i] = 0.5; }
    double eps0 = 8.854e-12;
    double epsf = dt*sigma/(2*eps0*epsr);
    for (int i = nx/2; i < nx; i++) ca[i] = (1 - epsf)/(1 + epsf);
    for (int i = nx/2; i < nx; i++) cb[i] = 0.5/(epsr*(1 + epsf));
}

/* ------------------------------------------------------- 2-D set-up */

/* Sullivan PML profiles in double, the PyCUDA host arithmetic of
 * fd2d/pycuda/fd2d_3_2.py:264-273 (xm = 0.33*((npml-n)/npml)**3 is a libm
 * pow(x, 3.0), as Python's float power), preceded by the default fills of the
 * allocation, fd2d/pycuda/fd2d_3_2.py:293-304 (f*1 = 0, everything else 1).
 * Arrays are in pmlayer order: fx1 fx2 fx3 (nx) fy1 fy2 fy3 (ny) gx2 gx3 (nx)
 * gy2 gy3 (ny).                                                            */
void orc_fd2d_pmlparam_f64(int nx, int ny, int npml,
                           double *fx1, double *fx2, double *fx3,
                           double *fy1, double *fy2, double *fy3,
                           double *gx2, double *gx3, double *gy2, double *gy3)
{
    for (int i = 0; i < nx; i++) { fx1[i] = 0.0; fx2[i] = 1.0; fx3[i] = 1.0; gx2[i] = 1.0; gx3[i] = 1.0; }
    for (int j = 0; j < ny; j++) { fy1[j] = 0.0; fy2[j] = 1.0; fy3[j] = 1.0; gy2[j] = 1.0; gy3[j] = 1.0; }
    for (int n = 0; n < npml; n++) {
        double xm = 0.33*pow((double)(npml-n)/npml, 3.0);
        double xn = 0.33*pow((npml-n-0.5)/npml, 3.0);
        fx1[n] = fx1[nx-2-n] = fy1[n] = fy1[ny-2-n] = xn;
        fx2[n] = fx2[nx-2-n] = fy2[n] = fy2[ny-2-n] = 1/(1+xn);
        gx2[n] = gx2[nx-1-n] = gy2[n] = gy2[ny-1-n] = 1/(1+xm);
        fx3[n] = fx3[nx-2-n] = fy3[n] = fy3[ny-2-n] = (1-xn)/(1+xn);
        gx3[n] = gx3[nx-1-n] = gy3[n] = gy3[ny-1-n] = (1-xm)/(1+xm);
    }
}

/* Lossy dielectric cylinders by 3x3 sub-cell sampling, in double.
 *
 * Single cylinder at (nx/2-1, ny/2-1), float sub-sample offsets m/3: the Numba
 * dielectric of fd2d/pycuda/fd2d_3_4.py:233-254 (without its fastmath), i.e.
 *     x = nx/2-1-i+m/3, y = ny/2-1-j+n/3, d = sqrt(x*x+y*y),
 *     d <= rgrid  =>  epsn += (epsr-1)/9, cond += sigma/9,
 *     naz = 1/(epsn + cond*dt/eps0), nbz = cond*dt/eps0,
 * over i, j in [npml, n-npml); other cells keep naz = 1, nbz = 0.
 * Several cylinders generalise this: a sub-sample takes the first cylinder c
 * (in list order) with hypot((cx[c]-i)+m/3, (cy[c]-j)+n/3) <= rad[c].
 * int_sub != 0 selects the C programs' integer sub-sample offsets m/3 == 0 and
 * integer centre arithmetic (fd2d/clang/fd2d_3_4.c:149-176), for which cx, cy
 * are ignored and every cylinder is centred at (nx/2-1, ny/2-1).
 * Outputs are double; the PyCUDA path stores them as float32.             */
void orc_fd2d_dielectric_f64(int nx, int ny, int npml, double dt,
                             int ncyl, const double *cx, const double *cy,
                             const double *rad, const double *epsr, const double *sigma,
                             int int_sub, double *naz, double *nbz)
{
    const double eps0 = 8.854e-12;
    for (size_t n = 0; n < (size_t)nx*ny; n++) { naz[n] = 1.0; nbz[n] = 0.0; }
    #pragma omp parallel for schedule(static)
    for (int i = npml; i < nx-npml; i++) {
        for (int j = npml; j < ny-npml; j++) {
            double epsn = 1.0;
            double cond = 0.0;
            for (int m = -1; m <= 1; m++) {
                for (int n = -1; n <= 1; n++) {
                    for (int c = 0; c < ncyl; c++) {
                        double x, y;
                        if (int_sub) {
                            x = nx/2-1-i+m/3;
                            y = ny/2-1-j+n/3;
                        } else {
                            x = cx[c] - i + m/3.0;
                            y = cy[c] - j + n/3.0;
                        }
                        double d = sqrt(x*x + y*y);
                        if (d <= rad[c]) {
                            epsn += (epsr[c] - 1)/9;
                            cond += sigma[c]/9;
                            break;
                        }
                    }
                }
            }
            naz[(size_t)i*ny+j] = 1/(epsn + cond*dt/eps0);
            nbz[(size_t)i*ny+j] = cond*dt/eps0;
        }
    }
}

/* --------------------------------------------------- precision instances */

#define REAL float
#define SUF f32
#define HALF 0.5f
#include "fdtd_oracle_impl.h"
#undef REAL
#undef SUF
#undef HALF

#define REAL double
#define SUF f64
#define HALF 0.5
#include "fdtd_oracle_impl.h"
#undef REAL
#undef SUF
#undef HALF
