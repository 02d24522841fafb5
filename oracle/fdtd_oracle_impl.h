/*
 * fdtd_oracle_impl.h -- precision-generic body of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY: included twice by fdtd_oracle.c, once with
 * REAL=float / SUF=f32 and once with REAL=double / SUF=f64.  Nothing in the
 * product (simulation_amd/, libfdtd_hip.so) includes, links or loads it.
 *
 * Every update restates the serial meaning of a reference function in the
 * same floating-point expression order (operands, association and the
 * literal 0.5 / 0.5f), so that an IEEE build without contraction
 * (-ffp-contract=off, no -ffast-math) reproduces the reference C programs
 * bit for bit.  The literal "HALF" is 0.5f in the float instance, as in the
 * reference's FP32 test_*.c programs, and 0.5 in the double instance, as in
 * its FP64 fd*_*.c programs.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUF)

/* ------------------------------------------------------------------ 1-D */

/* E update for i in [1, nx), then the source.
 * Forms, chosen by which coefficient arrays are given:
 *   ca == NULL, cb == NULL : ex[i] += 0.5*(hy[i-1]-hy[i])
 *                            fd1d/clang/test_1_2.c:18-24, fd1d/clang/fd1d_1_2.c:151-157
 *   ca == NULL, cb != NULL : ex[i] += cb[i]*(hy[i-1]-hy[i])
 *                            fd1d/clang/test_1_3.c (exfield), fd1d/pycuda/fd1d_1_3.py:46-53
 *   ca != NULL, cb != NULL : ex[i] = ca[i]*ex[i] + cb[i]*(hy[i-1]-hy[i])
 *                            fd1d/clang/fd1d_1_5.c:25-31, fd1d/pycuda/fd1d_1_5.py:49-56
 * Source: src_mode 0 none, 1 hard (ex[pos] = val, fd1d/clang/test_1_2.c:23),
 *         2 additive (ex[pos] += val, fd1d/clang/fd1d_1_5.c:30).            */
void FN(orc_fd1d_exfield)(int t, int nx, const REAL *ca, const REAL *cb,
                          REAL *ex, const REAL *hy,
                          int src_mode, int src_pos, REAL src_val)
{
    (void)t;
    if (!ca && !cb) {
        for (int i = 1; i < nx; i++)
            ex[i] += HALF * (hy[i-1] - hy[i]);
    } else if (!ca) {
        for (int i = 1; i < nx; i++)
            ex[i] += cb[i] * (hy[i-1] - hy[i]);
    } else {
        for (int i = 1; i < nx; i++)
            ex[i] = ca[i] * ex[i] + cb[i] * (hy[i-1] - hy[i]);
    }
    if (src_mode == 1) ex[src_pos] = src_val;
    else if (src_mode == 2) ex[src_pos] += src_val;
}

/* Two-step-history absorbing boundary, then the H update for i in [0, nx-1).
 * fd1d/clang/fd1d_1_2.c:160-167 (bc != NULL); fd1d/clang/fd1d_1_1.c hyfield
 * without the boundary (bc == NULL).                                       */
void FN(orc_fd1d_hyfield)(int nx, REAL *ex, REAL *hy, REAL *bc)
{
    if (bc) {
        ex[0] = bc[0]; bc[0] = bc[1]; bc[1] = ex[1];
        ex[nx-1] = bc[3]; bc[3] = bc[2]; bc[2] = ex[nx-2];
    }
    for (int i = 0; i < nx-1; i++)
        hy[i] += HALF * (ex[i] - ex[i+1]);
}

/* The 1-D driver loop of fd1d/clang/fd1d_1_2.c:180-183 / fd1d/pycuda/fd1d_1_2.py:87-89
 * for steps t = t0+1 .. t0+nsteps.  src_wave[s] is the source value of step
 * t0+1+s.  After each step, probe_out[p*nsteps+s] = ex[probe_k[p]].        */
void FN(orc_fd1d_run)(int nx, int t0, int nsteps, const REAL *ca, const REAL *cb,
                      REAL *ex, REAL *hy, REAL *bc,
                      int src_mode, int src_pos, const REAL *src_wave,
                      int nprobe, const int *probe_k, REAL *probe_out)
{
    for (int s = 0; s < nsteps; s++) {
        int t = t0 + 1 + s;
        FN(orc_fd1d_exfield)(t, nx, ca, cb, ex, hy, src_mode, src_pos,
                             src_wave ? src_wave[s] : (REAL)0);
        FN(orc_fd1d_hyfield)(nx, ex, hy, bc);
        for (int p = 0; p < nprobe; p++)
            probe_out[(size_t)p * nsteps + s] = ex[probe_k[p]];
    }
}

/* ------------------------------------------------------------------ 2-D */

/* D update for i in [1,nx), j in [1,ny).
 * pml given:  fd2d/clang/test_3_4.c:68-79 (same body fd2d/pycuda/fd2d_3_2.py:220-228)
 * pml NULL :  free-space form fd2d/clang/test_3_1.c (dfield), fd2d/clang/fd2d_3_1.c:18-29 */
void FN(orc_fd2d_dfield)(int nx, int ny, const REAL *gx2, const REAL *gx3,
                         const REAL *gy2, const REAL *gy3,
                         REAL *dz, const REAL *hx, const REAL *hy)
{
    if (gx2) {
        #pragma omp parallel for schedule(static)
        for (int i = 1; i < nx; i++) {
            for (int j = 1; j < ny; j++) {
                size_t n = (size_t)i*ny + j;
                dz[n] = gx3[i] * gy3[j] * dz[n] + gx2[i] * gy2[j] * HALF * (hy[n] - hy[n-ny] - hx[n] + hx[n-1]);
            }
        }
    } else {
        #pragma omp parallel for schedule(static)
        for (int i = 1; i < nx; i++) {
            for (int j = 1; j < ny; j++) {
                size_t n = (size_t)i*ny + j;
                dz[n] += HALF * (hy[n] - hy[n-ny] - hx[n] + hx[n-1]);
            }
        }
    }
}

/* E from D over all cells.
 * nbz given: fd2d/clang/test_3_4.c:92-102 (lossy form with the iz accumulator)
 * nbz NULL : fd2d/clang/fd2d_3_1.c:32-41 (ez = naz*dz)                       */
void FN(orc_fd2d_efield)(int nx, int ny, const REAL *naz, const REAL *nbz,
                         const REAL *dz, REAL *iz, REAL *ez)
{
    if (nbz) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; i++) {
            for (int j = 0; j < ny; j++) {
                size_t n = (size_t)i*ny + j;
                ez[n] = naz[n] * (dz[n] - iz[n]);
                iz[n] += nbz[n] * ez[n];
            }
        }
    } else {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nx; i++) {
            for (int j = 0; j < ny; j++) {
                size_t n = (size_t)i*ny + j;
                ez[n] = naz[n] * dz[n];
            }
        }
    }
}

/* H update for i in [0,nx-1), j in [0,ny-1).
 * pml given: fd2d/clang/test_3_4.c:114-126 (same body fd2d/pycuda/fd2d_3_2.py:249-260)
 * pml NULL : fd2d/clang/fd2d_3_1.c:44-54 (ihx, ihy untouched)                 */
void FN(orc_fd2d_hfield)(int nx, int ny, const REAL *fx1, const REAL *fx2, const REAL *fx3,
                         const REAL *fy1, const REAL *fy2, const REAL *fy3,
                         const REAL *ez, REAL *ihx, REAL *ihy, REAL *hx, REAL *hy)
{
    if (fx1) {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nx-1; i++) {
            for (int j = 0; j < ny-1; j++) {
                size_t n = (size_t)i*ny + j;
                ihx[n] += ez[n] - ez[n+1];
                ihy[n] += ez[n] - ez[n+ny];
                hx[n] = fy3[j] * hx[n] + fy2[j] * (HALF * ez[n] - HALF * ez[n+1] + fx1[i] * ihx[n]);
                hy[n] = fx3[i] * hy[n] - fx2[i] * (HALF * ez[n] - HALF * ez[n+ny] + fy1[j] * ihy[n]);
            }
        }
    } else {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < nx-1; i++) {
            for (int j = 0; j < ny-1; j++) {
                size_t n = (size_t)i*ny + j;
                hx[n] += HALF * (ez[n] - ez[n+1]);
                hy[n] -= HALF * (ez[n] - ez[n+ny]);
            }
        }
    }
}

/* The 2-D TM time loop in the serial order of fd2d/clang/fd2d_3_1.c:72-76
 * with the PML/lossy bodies above: for t = t0+1 .. t0+nsteps
 *   dfield; hard sources dz[si,sj] = src_wave[k*nsteps+s] (fd2d/clang/test_3_1.c:28,
 *   fd2d/pycuda/fd2d_3_1.py:69); efield; hfield;
 *   probe_out[p*nsteps+s] = ez[pi,pj].
 * pml10 points to the ten 1-D profiles in the order fx1 fx2 fx3 fy1 fy2 fy3
 * gx2 gx3 gy2 gy3 (the pmlayer order of fd2d/pycuda/fd2d_3_2.py:204-209),
 * or is NULL for the free-space form.                                       */
void FN(orc_fd2d_tm_run)(int nx, int ny, int t0, int nsteps,
                         const REAL *const *pml10, const REAL *naz, const REAL *nbz,
                         REAL *dz, REAL *iz, REAL *ez, REAL *hx, REAL *hy,
                         REAL *ihx, REAL *ihy,
                         int nsrc, const int *src_i, const int *src_j, const REAL *src_wave,
                         int nprobe, const int *probe_i, const int *probe_j, REAL *probe_out)
{
    const REAL *fx1 = 0, *fx2 = 0, *fx3 = 0, *fy1 = 0, *fy2 = 0, *fy3 = 0;
    const REAL *gx2 = 0, *gx3 = 0, *gy2 = 0, *gy3 = 0;
    if (pml10) {
        fx1 = pml10[0]; fx2 = pml10[1]; fx3 = pml10[2];
        fy1 = pml10[3]; fy2 = pml10[4]; fy3 = pml10[5];
        gx2 = pml10[6]; gx3 = pml10[7]; gy2 = pml10[8]; gy3 = pml10[9];
    }
    for (int s = 0; s < nsteps; s++) {
        FN(orc_fd2d_dfield)(nx, ny, gx2, gx3, gy2, gy3, dz, hx, hy);
        for (int k = 0; k < nsrc; k++)
            dz[(size_t)src_i[k]*ny + src_j[k]] = src_wave[(size_t)k*nsteps + s];
        FN(orc_fd2d_efield)(nx, ny, naz, nbz, dz, iz, ez);
        FN(orc_fd2d_hfield)(nx, ny, fx1, fx2, fx3, fy1, fy2, fy3, ez, ihx, ihy, hx, hy);
        for (int p = 0; p < nprobe; p++)
            probe_out[(size_t)p*nsteps + s] = ez[(size_t)probe_i[p]*ny + probe_j[p]];
    }
    (void)t0;
}

/* Sum of squares of the three TM fields, accumulated in double in row-major
 * order.  Not a reference function: the reference has no energy accounting;
 * this is the checker of the device reduction.                              */
double FN(orc_fd2d_energy)(int nx, int ny, const REAL *ez, const REAL *hx, const REAL *hy)
{
    double acc = 0.0;
    for (size_t n = 0; n < (size_t)nx*ny; n++) {
        double e = (double)ez[n], a = (double)hx[n], b = (double)hy[n];
        acc += e*e + a*a + b*b;
    }
    return acc;
}

#undef FN
#undef CAT
#undef CAT2
