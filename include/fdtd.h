/*
 * fdtd.h -- C-ABI of libfdtd_hip.so, the MI355X (gfx950) FDTD hot path.
 *
 * Drop-in boundary for the PyCUDA programs of dsarvan/simulation
 * (fd1d/pycuda/*.py, fd2d/pycuda/*.py).  There, the "operator API" is
 * `SourceModule(kernel).get_function(name)` and a call
 * `f(*args, grid=, block=)` on gpuarrays (fd2d/pycuda/fd2d_3_2.py:333-341).
 * Here each such kernel is an extern "C" entry point with the same argument
 * order and meaning, taking raw device pointers (e.g. torch.Tensor.data_ptr())
 * and a HIP stream (hipStream_t passed as void*, NULL = the null stream).
 *
 * Conventions
 *   - Return value: 0 on success; FDTD_EINVAL (< 0) when an argument is
 *     rejected before any device work is queued; a positive hipError_t when
 *     the HIP runtime reports an error.  fdtd_last_error() gives a message.
 *     The reference checks no error at all (its CUDA errors surface at the
 *     next drv.Context.synchronize(), fd2d/pycuda/fd2d_3_2.py:343).
 *   - Launches are asynchronous on `stream`, in order, like PyCUDA's calls
 *     on the null stream.  Entry points that return a host result say so.
 *   - No entry point allocates device memory: every buffer, including the
 *     ping-pong partner set and the reduction work area, is the caller's.
 *   - 2-D arrays are row-major nx*ny with n = i*ny + j, j contiguous
 *     (fd2d/pycuda/fd2d_3_1.py:63, :106-111).
 *   - Field arithmetic is IEEE FP32 without contraction, in the expression
 *     order of the reference kernels, so results equal the reference's FP32
 *     C programs bit for bit.  FP64 is used for coefficient set-up and
 *     energy accounting only.
 *   - Sources are evaluated by the caller (host) and passed as values; the
 *     cell that owns a source applies it, so the reference's source/boundary
 *     races (a thread writing a cell another thread updates) cannot occur.
 */
#ifndef FDTD_H
#define FDTD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD_ABI_VERSION 4

#define FDTD_OK 0
#define FDTD_EINVAL (-1)

#define FDTD_MAX_SOURCES 16
#define FDTD_MAX_PROBES 16
#define FDTD_MAX_CYLINDERS 64
#define FDTD_ENERGY_WORK 1025   /* doubles in the energy work area */

/* Source modes of the 1-D E update. */
#define FDTD_SRC_NONE 0
#define FDTD_SRC_HARD 1        /* ex[pos] = value   (fd1d/pycuda/fd1d_1_2.py:49) */
#define FDTD_SRC_ADD 2         /* ex[pos] += value  (fd1d/pycuda/fd1d_1_5.py:56) */

/* The PML profile set.  Replaces the device struct `pmlayer` and its device
 * pointer table (fd2d/pycuda/fd2d_3_2.py:204-209, :306-317): here a host
 * struct of device pointers, copied into kernel arguments.  fx*, gx* have nx
 * elements, fy*, gy* have ny.                                               */
typedef struct fdtd_pml {
    float *fx1, *fx2, *fx3;
    float *fy1, *fy2, *fy3;
    float *gx2, *gx3;
    float *gy2, *gy3;
} fdtd_pml;

/* Replaces the device struct `medium` (fd2d/pycuda/fd2d_3_4.py:96-98, :322-324). */
typedef struct fdtd_medium {
    float *naz, *nbz;
} fdtd_medium;

/* ------------------------------------------------------------ library */

int fdtd_abi_version(void);
/* Message of the last failed call on this host thread ("" if none). */
const char *fdtd_last_error(void);
/* Loads every device code object of the library on the current device.  The
 * HIP runtime otherwise loads a translation unit's code at the first launch
 * of one of its kernels, so that first launch would carry the load.  No
 * counterpart in the reference (PyCUDA compiles and loads its module up
 * front, SourceModule at fd2d/pycuda/fd2d_3_4.py:353).  Optional; callable
 * any number of times.                                                      */
int fdtd_preload(void);

/* ---------------------------------------------------- host waveforms
 * Evaluated on the host with the host libm, in the reference's expression
 * order, so values equal its C programs':
 *   gaussian   fd2d/pycuda/fd2d_3_1.py:52-55, fd1d/clang/test_1_2.c:13-15
 *   sinusoidal fd2d/pycuda/fd2d_3_2.py:212-216, fd2d/clang/test_3_2.c:22-25 */
float fdtd_gaussian_f32(int t, int t0, float sigma);
double fdtd_gaussian_f64(int t, int t0, double sigma);
float fdtd_sinusoidal_f32(int t, float ds, float freq);
double fdtd_sinusoidal_f64(int t, double ds, double freq);

/* ------------------------------------------------------------- 1-D */

/* E update for i in [1,nx) then the source.  Replaces `exfield` of
 * fd1d/pycuda/fd1d_1_5.py:49-56 (ca, cb), fd1d/pycuda/fd1d_1_3.py:46-53
 * (ca NULL: ca = 1) and fd1d/pycuda/fd1d_1_2.py:43-50 (ca, cb NULL: cb = 0.5).
 * src_mode FDTD_SRC_*, src_val evaluated by the caller for step t.          */
int fd1d_exfield(int t, int nx, const float *ca, const float *cb,
                 float *ex, const float *hy,
                 int src_mode, int src_pos, float src_val, void *stream);
int fd1d_exfield_f64(int t, int nx, const double *ca, const double *cb,
                     double *ex, const double *hy,
                     int src_mode, int src_pos, double src_val, void *stream);

/* Two-step absorbing boundary (bc[4] state) then H update for i in [0,nx-1).
 * Replaces `hyfield` of fd1d/pycuda/fd1d_1_2.py:53-61; bc NULL gives the
 * boundary-free form of fd1d/pycuda/fd1d_1_1.py:53-58.  nx >= 3.            */
int fd1d_hyfield(int nx, float *ex, float *hy, float *bc, void *stream);
int fd1d_hyfield_f64(int nx, double *ex, double *hy, double *bc, void *stream);

/* nsteps steps of exfield+hyfield in ONE launch (single workgroup, fields in
 * LDS), for t = t0+1 .. t0+nsteps.  Replaces the driver loop of
 * fd1d/pycuda/fd1d_1_2.py:87-89 (2 launches per step there).
 * src_wave: device, src_wave[s] = source value of step t0+1+s (NULL: none).
 * probe_k: host array of nprobe cell indices; probe_out: device,
 * probe_out[p*nsteps + s] = ex[probe_k[p]] after step t0+1+s.
 * Requires nx <= fd1d_run_max_nx(sizeof element).                           */
int fd1d_run(int nx, int t0, int nsteps, const float *ca, const float *cb,
             float *ex, float *hy, float *bc,
             int src_mode, int src_pos, const float *src_wave,
             int nprobe, const int *probe_k, float *probe_out, void *stream);
int fd1d_run_f64(int nx, int t0, int nsteps, const double *ca, const double *cb,
                 double *ex, double *hy, double *bc,
                 int src_mode, int src_pos, const double *src_wave,
                 int nprobe, const int *probe_k, double *probe_out, void *stream);
int fd1d_run_max_nx(int elem_size);

/* ca, cb of a lossy half space i >= nx/2, computed on the device in FP64 and
 * stored FP32, as fd1d/pycuda/fd1d_1_5.py:71-78 (host FP64 there).          */
int fd1d_dielectric(int nx, double dt, double epsr, double sigma,
                    float *ca, float *cb, void *stream);
/* The same coefficients stored FP64, for the FP64 runs of fd1d/clang/fd1d_1_5.c:44-55. */
int fd1d_dielectric_f64(int nx, double dt, double epsr, double sigma,
                        double *ca, double *cb, void *stream);

/* ------------------------------------------------------------- 2-D */

/* D update for i in [1,nx), j in [1,ny), then hard sources.  Replaces
 * `dfield(int t,int nx,int ny,pmlayer *pml,float *dz,float *hx,float *hy)`
 * of fd2d/pycuda/fd2d_3_2.py:219-233; pml NULL gives the free-space form of
 * fd2d/pycuda/fd2d_3_1.py:58-70.  Sources: nsrc <= FDTD_MAX_SOURCES host
 * arrays of cell (src_i, src_j) and value; dz[cell] = value after the update
 * (later sources win), applied by the thread that owns the cell.            */
int fd2d_dfield(int t, int nx, int ny, const fdtd_pml *pml,
                float *dz, const float *hx, const float *hy,
                int nsrc, const int *src_i, const int *src_j, const float *src_val,
                void *stream);

/* E from D over all cells.  Replaces `efield(int nx,int ny,medium *md,float
 * *dz,float *iz,float *ez)` of fd2d/pycuda/fd2d_3_4.py:175-185; md->nbz NULL
 * gives `ez = naz*dz` of fd2d/pycuda/fd2d_3_1.py:73-82 (iz unused).          */
int fd2d_efield(int nx, int ny, const fdtd_medium *md,
                const float *dz, float *iz, float *ez, void *stream);

/* H update for i in [0,nx-1), j in [0,ny-1).  Replaces `hfield(int nx,int
 * ny,pmlayer *pml,float *ez,float *ihx,float *ihy,float *hx,float *hy)` of
 * fd2d/pycuda/fd2d_3_2.py:248-260; pml NULL gives the free-space form of
 * fd2d/pycuda/fd2d_3_1.py:85-95 (ihx, ihy unused).                          */
int fd2d_hfield(int nx, int ny, const fdtd_pml *pml, const float *ez,
                float *ihx, float *ihy, float *hx, float *hy, void *stream);

/* Sullivan PML profiles computed on the device in FP64, stored FP32, over the
 * default fills (f*1 = 0, others 1).  Replaces `pmlparam` of
 * fd2d/pycuda/fd2d_3_2.py:264-273 (host FP64, 20*npml scalar copies there)
 * and the fills of :293-304.  Requires 0 <= npml <= min(nx,ny) - 2.        */
int fd2d_pmlparam(int nx, int ny, int npml, const fdtd_pml *pml, void *stream);

/* Lossy dielectric cylinders by 3x3 sub-cell sampling, on the device in FP64,
 * stored FP32.  Replaces the Numba `dielectric` of
 * fd2d/pycuda/fd2d_3_4.py:233-254 (one cylinder centred at (nx/2-1, ny/2-1):
 * cyl = {nx/2.0-1, ny/2.0-1, rgrid, epsr, sigma}).  cyl: host array of ncyl
 * records {cx, cy, radius, epsr, sigma}; a sub-sample takes the first
 * cylinder that contains it; cells outside [npml, n-npml) get naz=1, nbz=0.
 * int_sub != 0 selects the integer sub-sampling of fd2d/clang/fd2d_3_4.c:
 * 149-176 (cx, cy ignored).                                                 */
int fd2d_dielectric(int nx, int ny, int npml, double dt,
                    int ncyl, const double *cyl, int int_sub,
                    float *naz, float *nbz, void *stream);

/* Sum over the window [i0,i1) x [j0,j1) of ez^2 + hx^2 + hy^2, accumulated
 * in FP64 in a fixed order (deterministic).  work: device, FDTD_ENERGY_WORK
 * doubles; the result is left in work[0].  If result != NULL the call
 * synchronises `stream` and stores the value there.  Not a reference
 * function: the reference has no energy or instability accounting.         */
int fd2d_energy(int nx, int ny, int i0, int i1, int j0, int j1,
                const float *ez, const float *hx, const float *hy,
                double *work, double *result, void *stream);

/* fd2d_energy without a work area, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The partial sums live in a library-owned
 * static device array of FDTD_ENERGY_WORK doubles (8200 B per GPU, held until
 * the process ends); the reduction needs exactly that for any window, so no
 * window exceeds it.  result: device pointer to one double, written
 * asynchronously on `stream` (the call neither synchronises nor copies to
 * the host).  Allocates nothing.  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other host threads)
 * race.  Nothing in the array needs to start at zero.                        */
int fd2d_energy_noalloc(int nx, int ny, int i0, int i1, int j0, int j1,
                        const float *ez, const float *hx, const float *hy,
                        double *result, void *stream);

/* FP64 energy of the difference of two field sets over equal wn x wm
 * windows -- window (ia0, ja0) of grid a (nxa x nya) against (ib0, jb0) of
 * grid b: sum (ez_a - ez_b)^2 + (hx_a - hx_b)^2 + (hy_a - hy_b)^2.  Used to
 * measure reflected energy against a reference run on a larger grid.  No
 * reference counterpart.  Same work buffer and result handling as
 * fd2d_energy.                                                              */
int fd2d_energy_diff(int wn, int wm, int nxa, int nya, int ia0, int ja0, const float *eza,
                     const float *hxa, const float *hya, int nxb, int nyb, int ib0, int jb0,
                     const float *ezb, const float *hxb, const float *hyb, double *work,
                     double *result, void *stream);

/* Fused, temporally blocked TM time stepping.  One call advances the state
 * by nsteps leapfrog steps t = t0+1 .. t0+nsteps with the serial semantics
 * dfield -> hard sources -> efield -> hfield of fd2d/pycuda/fd2d_3_4.py:
 * 364-373 (config-2 form, fd2d_3_2.py dfield/hfield + fd2d_3_4.py efield),
 * i.e. the same values as that many fd2d_dfield/efield/hfield calls.
 * Each launch fuses `depth` steps; launches alternate between the caller's
 * state set and the partner set (dz2 ...), and the call always ends with the
 * result in the caller's set.  ez is written for the final step only.
 *   src_wave: device [nsrc][wave_len]; the value of step t for source k is
 *             src_wave[k*wave_len + t-1] (t0+nsteps <= wave_len).
 *   probe_out: device [nprobe][probe_len]; ez at probe p after step t goes
 *             to probe_out[p*probe_len + t-1] (t0+nsteps <= probe_len).
 *   md.nbz NULL: lossless (ez = naz*dz, iz untouched); pml all NULL: free
 *   space (ihx, ihy untouched).
 *   depth/rows/cols: temporal depth, rows per block and columns per lane;
 *   0 selects the built-in defaults (see fd2d_tm_plan).
 *   form: FDTD_TM_FORM_AUTO (0) lets the library choose (currently the
 *   single-wave strips, the faster form on MI355X); FDTD_TM_FORM_STRIP (1)
 *   sweeps everything in single-wave strips (any cols);
 *   FDTD_TM_FORM_MULTIWAVE (2) sweeps the interior of the band in workgroups
 *   of four waves that exchange their edge columns through LDS (one column
 *   per lane; `rows` = rows of such a block) and the rest in single-wave
 *   strips.  All forms give the same bits.
 *   band_rows/band_cols: the caller's promise that row indices
 *   [band_rows[0], band_rows[1]) hold the default profile values
 *   (fx1 = 0, fx2 = fx3 = gx2 = gx3 = 1) and column indices
 *   [band_cols[0], band_cols[1]) likewise for fy*, gy* (as fd2d_pmlparam
 *   leaves [npml, n-1-npml)).  Blocks of the grid wholly inside both bands
 *   run a body with those values as constants (same results, fewer loads
 *   and multiplies by one).  An empty range promises nothing.              */
typedef struct fd2d_tm_desc {
    int nx, ny;
    float *dz, *iz, *ez, *hx, *hy, *ihx, *ihy;
    float *dz2, *iz2, *hx2, *hy2, *ihx2, *ihy2;
    fdtd_medium md;
    fdtd_pml pml;
    int nsrc;
    int src_i[FDTD_MAX_SOURCES], src_j[FDTD_MAX_SOURCES];
    const float *src_wave;
    int wave_len;
    int nprobe;
    int probe_i[FDTD_MAX_PROBES], probe_j[FDTD_MAX_PROBES];
    float *probe_out;
    int probe_len;
    int depth, rows, cols;
    int band_rows[2], band_cols[2];
    int form;                     /* FDTD_TM_FORM_* */
    /* Optional workspace of tile_elems >= fd2d_tm_tile_elems(d, nsteps)
     * unsigned ints (NULL: none).  With it, the first fused launch of a
     * call records which interior tiles read only naz = 1, nbz = +0 and
     * iz = +0 (free space at rest), and the call's later launches neither
     * read naz, nbz, iz there nor store iz (it stays +0 in both state sets:
     * iz + 0 * ez is +0; every result bit is the same as without the
     * workspace, for finite fields).
     * tile_keep = 1 promises that naz and nbz hold what they held when the
     * flags in tile_work were last recorded, and that iz and iz2 were
     * written by nothing but those runs since: a call with the same arrays,
     * grid, rows, columns and a depth no deeper than the recording one then
     * uses those flags from its first launch on (new arrays that reuse the
     * memory of freed ones break that promise: pass tile_keep = 0 on the
     * first call of a new medium or state).  tile_keep = 0: flags of
     * earlier calls are never used.                                        */
    unsigned *tile_work;
    size_t tile_elems;
    int tile_keep;
} fd2d_tm_desc;

#define FDTD_TM_FORM_AUTO 0
#define FDTD_TM_FORM_STRIP 1
#define FDTD_TM_FORM_MULTIWAVE 2

int fd2d_tm_run(const fd2d_tm_desc *d, int t0, int nsteps, void *stream);

/* The launch plan fd2d_tm_run would use for this descriptor and nsteps:
 * number of launches and the chosen depth/rows/cols.  Defaults (measured on
 * MI355X): one column per lane; depth 10 from nx = 2048 on, else 8; rows =
 * nx/128 for 4096 <= nx < 12288, else nx/170, clamped to [8, 512]; the
 * multi-wave form: depth 8, rows = nx/32 clamped to [32, 512].  Depths go up
 * to 12.  Host only.                                                       */
int fd2d_tm_plan(const fd2d_tm_desc *d, int nsteps,
                 int *launches, int *depth, int *rows, int *cols);

/* Elements of fd2d_tm_desc.tile_work the plan of (d, nsteps) uses: one per
 * tile of the single-wave strip form (0 for the multi-wave form, which does
 * not use it).  Host only.  No reference counterpart.                      */
size_t fd2d_tm_tile_elems(const fd2d_tm_desc *d, int nsteps);

/* ------------------------------------------------------ batched grids
 * Many independent TM grids of different sizes advanced together by one
 * persistent kernel per fused step group.  No reference counterpart: the
 * reference runs one grid per program (fd2d/pycuda/fd2d_3_4.py:269-389);
 * each grid here evolves exactly as fd2d_tm_run would evolve it alone (bit
 * for bit).  All grids of a batch share the form (lossy PML: md.nbz and the
 * ten profiles given) and advance by the same steps.                        */
typedef struct fd2d_grid {
    int nx, ny;
    float *dz, *iz, *ez, *hx, *hy, *ihx, *ihy;       /* state (ez: final step only) */
    float *dz2, *iz2, *hx2, *hy2, *ihx2, *ihy2;     /* partner state set */
    fdtd_medium md;
    fdtd_pml pml;
    int band_rows[2], band_cols[2];                 /* as in fd2d_tm_desc */
    int src_i, src_j;                               /* hard source; src_i < 0: none */
    const float *src_wave;                          /* value of step t at [t-1] */
    int wave_len;
    int probe_i, probe_j;                           /* ez probe; probe_i < 0: none */
    float *probe_out;                               /* ez after step t at [t-1] */
    int probe_len;
} fd2d_grid;

typedef struct fd2d_batch_desc {
    int ngrids;
    const fd2d_grid *grids_dev;   /* device copy of the grid records */
    const int *units_dev;         /* device work list, 4 ints per unit */
    int nunits;
    int depth, rows, cols;        /* fixed when the work list is made (rows 0:
                                     the default, see fd2d_batch_units) */
    int waves;                    /* workgroups of the persistent launch */
    int vec;                      /* 1: every grid allows vector loads */
    int wave_len;                 /* smallest wave_len over grids with a source
                                     (INT_MAX: none), set by fd2d_batch_upload */
    int probe_len;                /* smallest probe_len over grids with a probe
                                     (INT_MAX: none), set by fd2d_batch_upload */
    unsigned *unit_work;          /* optional workspace of nunits unsigned ints
                                     (NULL: none; fd2d_batch_upload sets NULL).
                                     With it, the first launch of a run (or of
                                     each graph launch) records which interior
                                     units read only naz = 1 and nbz = 0, and
                                     the later launches skip reading naz and
                                     nbz there; same results.  The flags are
                                     recorded afresh by every run and graph
                                     launch, so a medium may change between
                                     them. */
} fd2d_batch_desc;

/* Number of work units (and so ints/4 of units_dev) a batch needs for the
 * given depth (0 = 8), rows per block (0 = 128, or the whole grid when it
 * has fewer rows) and columns per lane (0 = 1).  A unit is 4 ints: grid, strip,
 * first row, one past the last row.                                       */
int fd2d_batch_units(const fd2d_grid *grids, int ngrids, int depth, int rows, int cols,
                     int *nunits);

/* Validate the grids (host array), write the work list and the grid records
 * into the caller's device buffers (asynchronously on stream; the host
 * arrays may be released on return) and fill *out.  units_dev holds
 * 4*max_units ints, grids_dev ngrids records.  waves = 0 picks the number of
 * waves the device keeps resident.  The units are ordered for the waves of
 * the persistent launch (wave w takes units w, w + waves, ...): sorted by
 * estimated cost and dealt out round by round in alternating directions.   */
int fd2d_batch_upload(const fd2d_grid *grids, int ngrids, int depth, int rows, int cols,
                      int waves, fd2d_grid *grids_dev, int *units_dev, int max_units,
                      fd2d_batch_desc *out, void *stream);

/* Advance every grid by nsteps (t = t0+1 .. t0+nsteps).  t0+nsteps is
 * checked against the smallest wave_len / probe_len of the grids recorded in
 * *b by fd2d_batch_upload, and against the caller's max_wave_len /
 * max_probe_len (pass INT_MAX to rely on the recorded ones alone).  The
 * steps are split into an even number of launches of at most `depth` steps,
 * so that the result lands in the grids' own state sets: nsteps = 1, or an
 * odd nsteps with depth 1, is rejected.                                    */
int fd2d_batch_run(const fd2d_batch_desc *b, int t0, int nsteps, int max_wave_len,
                   int max_probe_len, void *stream);

/* The same launches captured into one executable HIP graph (returned as an
 * opaque handle), launched by fdtd_graph_launch and freed by
 * fdtd_graph_destroy.  stream must not be the null stream.                */
int fd2d_batch_graph(const fd2d_batch_desc *b, int t0, int nsteps, int max_wave_len,
                     int max_probe_len, void *stream, void **graph_exec);
int fdtd_graph_launch(void *graph_exec, void *stream);
int fdtd_graph_destroy(void *graph_exec);

/* ------------------------------------------------ learned absorbing boundary
 * No reference counterpart: the reference absorbs with the two-step 1-D ABC
 * (fd1d/pycuda/fd1d_1_2.py:53-61) and the Sullivan PML
 * (fd2d/pycuda/fd2d_3_2.py:126-135); this replaces a thick PML with a small
 * network applied at the outer three cells of every edge.
 *
 * Network: 108 features (the last four canonical boundary samples of a
 * position, newest first, 27 features each) -> 3 outputs.  With s = max|x|,
 * xf = x * (1/s) in FP32 and xh = bf16(xf):
 *   ez_next[d] = base[d] + s * y[d],  d = 0..2,
 *   y = L xf + (f(xh) - f(-xh)) / 2,
 *   f(x) = W3 bf16(gelu(W2 bf16(gelu(W1 x + b1)) + b2)) + b3,
 *   gelu(z) = z sigma(1.702 z),
 * base = the two-step absorbing boundary of fd1d/pycuda/fd1d_1_2.py:53-61 on
 * the three outer layers (ez_d(t+1) = ez_{d+1}(t-1)): a zero network is that
 * classical boundary.  The linear path L (FP32, fitted by ridge least
 * squares: lbc_gram, lbc_fit_linear) carries most of the correction; the
 * hidden layers (bf16 weights and activations, FP32 accumulation on MFMA,
 * trained by lbc_train_grad / lbc_adam) fit what it leaves.
 * Parameters: one FP32 vector of lbc_param_count() values, in the order
 * W1 (64x108, row-major), b1 (64), W2 (64x64), b2, W3 (3x64), b3, L (3x108);
 * the hidden layers read bf16 copies from a buffer of lbc_wbuf_elems()
 * uint16 values written by lbc_pack (and kept current by lbc_adam); its last
 * elements hold a flag, non-zero when W3 or b3 has a non-zero value: with a
 * zero output layer the hidden layers add exactly nothing and fd2d_lbc_run
 * skips them.
 *
 * Canonical sample (30 floats): 27 features at index (f*3 + d)*3 + o + 1 for
 * field f, feature depth d = 0..2 (grid depth D = d + 3: the three cells just
 * inside the predicted ones, so the network never reads its own output),
 * tangential offset o = -1..1 (tangential index clamped to the edge); f = 0
 * ez, 1 normal-coupled H, 2 tangential-coupled H:
 *   top    (i = D):      ez[D,t],      hy[D,t],       hx[D,t]
 *   bottom (i = nx-1-D): ez[nx-1-D,t], -hy[nx-2-D,t], hx[nx-1-D,t]
 *   left   (j = D):      ez[t,D],      -hx[t,D],      -hy[t,D]
 *   right  (j = ny-1-D): ez[t,ny-1-D], hx[t,ny-2-D],  -hy[t,ny-1-D]
 * with the tangential-coupled value 0 at offset +1 (it is staggered half a
 * cell along the edge; offsets -1 and 0 are the two points beside the cell)
 * and H in the last row or column of the window read as 0 (the H update
 * never runs there); then ez of grid depths 2, 1 and 0 at offset 0 (targets
 * and baseline only).  Positions: top then bottom (t in [0, ny), owning the
 * corners), then left and right (t in [3, nx-3)):
 * lbc_npos = 2 ny + 2 (nx - 6).                                             */
int lbc_param_count(void);
int lbc_wbuf_elems(void);
size_t lbc_train_smem(void);

/* Canonical samples of an nx x ny window (row stride ld) of ez/hx/hy:
 * out[lbc_npos][30].                                                        */
int lbc_sample(int nx, int ny, int ld, const float *ez, const float *hx, const float *hy,
               float *out, void *stream);

/* bf16 operand buffer from the FP32 parameters.                             */
int lbc_pack(const float *params, uint16_t *wbuf, void *stream);

/* Rows of features given as samples in device memory: row r's sample k
 * (k = 0 newest .. 3; its first 27 floats are read) starts at
 * src + offs[r] - k*tstride.  Writes the
 * normalised outputs y[r][3] and (if s is not NULL) the scales s[r].        */
int lbc_forward(int n, const float *src, const long long *offs, long long tstride,
                const float *params, const uint16_t *wbuf, float *y, float *s, void *stream);

/* Training gradient of  sum_r w_r |y_r - (T_r - base_r)/s_r|^2 / norm
 * (w_r = 0 for all-zero rows), targets T_r = tgt[r][3] or, with tgt NULL,
 * the next sample's ez of grid depths 0..2 (src + offs[r] + tstride); the
 * baseline reads the sample of age 1, which must be a full 30-float sample.
 * The linear path is fixed here: the hidden layers fit what it leaves, and
 * the gradient entries of L are zero.
 * nblocks workgroups each write one partial gradient (lbc_param_count
 * floats) to part[block] and its loss sum to part_loss[block]; lbc_adam
 * reduces them in a fixed order (deterministic).  noise > 0 adds uniform
 * noise of that amplitude to the normalised inputs (a hash of row, feature
 * and seed: reproducible), for a network that tolerates its own errors when
 * its outputs feed its later inputs.                                        */
int lbc_train_grad(int n, const float *src, const long long *offs, long long tstride,
                   const float *tgt, float norm, const float *params, const uint16_t *wbuf,
                   int nblocks, float *part, float *part_loss, float noise, unsigned seed,
                   void *stream);

/* Time stepping with the learned boundary.  The TM descriptor gives the grid
 * without a PML (pml fields NULL; md.nbz optional: lossy or lossless), the
 * partner state set, sources and probes (all at least fd2d_lbc_margin()
 * cells from every edge, and 3 depth + 6 when that is more: the frame
 * kernel's deepest region) and the launch shape of the blocked kernel
 * (depth at most 10, default 8).  Each launch of k <= depth steps is the blocked launch with PEC
 * edges followed by a frame kernel that recomputes the outer k+3 cells with
 * the network applied every step (the blocked launch is exact deeper in).
 * hist, hist2: [lbc_hist_elems(nx, ny)] floats each (npos x 4 x 30), the last four boundary
 * samples of every position (zero for fields at rest); the run leaves them
 * current in hist.  work: lbc_work_bytes(nx, ny) bytes of device memory for
 * the frame kernel's geometry tables and the interior tiles' free-space
 * flags (recorded by the first launch of every call, as with
 * fd2d_tm_desc.tile_work; rewritten by every call; one run at a time per
 * work buffer).  nsteps >= 2 (an even number of launches).                 */
#define FDTD_LBC_MARGIN 32
typedef struct fd2d_lbc_desc {
    const float *params;        /* lbc_param_count() FP32 parameters         */
    const uint16_t *wbuf;       /* bf16 operands (lbc_pack)                  */
    float *hist, *hist2;        /* boundary-sample history, ping-pong pair   */
    void *work;                 /* lbc_work_bytes(nx, ny) bytes               */
} fd2d_lbc_desc;
size_t lbc_hist_elems(int nx, int ny);
size_t lbc_work_bytes(int nx, int ny);
int fd2d_lbc_margin(void);
int fd2d_lbc_run(const fd2d_tm_desc *d, const fd2d_lbc_desc *b, int t0, int nsteps,
                 void *stream);

/* Reduces the partials; writes the gradient to grad_out (if not NULL); with
 * lr != 0 applies one Adam step (bias corrections for `step` >= 1) to the
 * parameters [first, first + count) (count 0: all), their m, v, and
 * refreshes their bf16 copies in wbuf.                                      */
int lbc_adam(int nblocks, const float *part, float *params, float *m, float *v, int step,
             float lr, float beta1, float beta2, float eps, uint16_t *wbuf, float *grad_out,
             int first, int count, void *stream);

/* ------------------------------------------------- running DFT planes
 * Replaces the device struct `ftrans` (fd2d/pycuda/fd2d_3_4.py:100-103,
 * fd1d/pycuda/fd1d_2_3.py:74-77): r_pt/i_pt are nf planes of the grid
 * (plane n at n * cells), r_in/i_in the nf transforms of the source probe. */
#define FDTD_MAX_FREQ 16
typedef struct fdtd_ftrans {
    float *r_pt, *i_pt;
    float *r_in, *i_in;
} fdtd_ftrans;

/* ------------------------------------- 1-D D/E form, lossy and Debye media
 * fd1d/pycuda/fd1d_2_1.py .. fd1d_2_3.py: D update with an additive source,
 * E from D through the medium (nax, nbx: lossy; with ncx, ndx: Debye
 * relaxation sx), the running DFT of ex, then fd1d_hyfield (two-step
 * boundary, H update).  Serial semantics of fd1d/clang/test_2_3.c:26-90,
 * FP32, bit-exact.                                                          */
/* Replaces the device struct `medium` of fd1d/pycuda/fd1d_2_3.py:68-71.    */
typedef struct fdtd_medium1d {
    float *nax, *nbx;
    float *ncx, *ndx;           /* NULL: the lossy form of fd1d_2_2 (no sx) */
} fdtd_medium1d;

/* `dxfield(t, nx, dx, hy)` (fd1d/pycuda/fd1d_2_3.py:96-102): dx update for
 * i in [1, nx), then dx[src_pos] += src_val (src_pos < 0: no source; the
 * reference adds gaussian(t, 50, 10) at 1).                                 */
int fd1d_dxfield(int t, int nx, float *dx, const float *hy, int src_pos, float src_val,
                 void *stream);
/* `exfield(nx, md, dx, ix, sx, ex)` (fd1d/pycuda/fd1d_2_3.py:105-113; with
 * md->ncx NULL, fd1d_2_2.py's lossy form; sx unused).                      */
int fd1d_exfield_dform(int nx, const fdtd_medium1d *md, const float *dx, float *ix, float *sx,
                       float *ex, void *stream);
/* `fourier(t, nf, nx, dt, freq, ex, ft)` (fd1d/pycuda/fd1d_2_3.py:79-93):
 * planes r_pt/i_pt of nx, probe ex[10]; freq host, twiddles as the 2-D
 * fd2d_fourier.                                                            */
int fd1d_fourier(int t, int nf, int nx, float dt, const float *freq, const float *ex,
                 const fdtd_ftrans *ft, void *stream);
/* `dielectric(nx, dt, chi, tau, epsr, sigma)` of fd1d/clang/test_2_3.c:75-90
 * (md->ncx NULL: test_2_2.c's lossy medium): free space below nx/2.        */
int fd1d_dielectric_dform(int nx, float dt, float chi, float tau, float epsr, float sigma,
                          const fdtd_medium1d *md, void *stream);

/* A whole 1-D D/E-form run, per step t = t0+1 .. t0+nsteps: dxfield with
 * dx[src_pos] += src_wave[t-1], exfield, fourier (nf > 0), fd1d_hyfield,
 * probes ex[probe_k[p]] -> probe_out[p*probe_len + t-1].                    */
typedef struct fd1d_dform_desc {
    int nx;
    float *dx, *ix, *sx, *ex, *hy, *bc;
    fdtd_medium1d md;
    int src_pos;                /* < 0: no source                           */
    const float *src_wave;      /* host                                     */
    int wave_len;
    int nf;
    const float *freq;          /* host                                     */
    float dt;
    fdtd_ftrans ft;
    int nprobe;
    int probe_k[FDTD_MAX_PROBES];
    float *probe_out;
    int probe_len;
    float *work;                /* NULL: the per-step kernels; else fused     */
    size_t work_elems;          /* >= fd1d_dform_work_elems(nx)               */
} fd1d_dform_desc;
/* With work, the run is fused: workgroups of 512 cells advance up to 32
 * steps per launch in LDS and registers (nf <= 3), reading one state set and
 * writing the other (work holds the partner set of dx ix sx ex hy bc), with
 * the same bits as the per-step kernels.                                    */
size_t fd1d_dform_work_elems(int nx);
int fd1d_dform_run(const fd1d_dform_desc *d, int t0, int nsteps, void *stream);

/* ------------------------------------------ plane wave (TF/SF) and DFT
 * The incident plane wave of fd2d/pycuda/fd2d_3_3.py:80-169 and
 * fd2d_3_4.py:121-226: a 1-D grid along y (ezi, hxi: ny cells) with the
 * two-step boundary (bc[4]) at both ends, and the total-field/scattered-
 * field corrections on the box of rows and columns npml-1 .. n-npml.
 * Serial semantics of fd2d/clang/test_3_4.c:38-147, FP32, bit-exact.       */

/* `ezinct(ny, ezi, hxi, bc)` (fd2d_3_4.py:140-148): ezi update, then the
 * boundary at both ends (one workgroup: every update precedes them).        */
int fd2d_ezinct(int ny, float *ezi, const float *hxi, float *bc, void *stream);
/* `dfield(t, nx, ny, pml, ezi, dz, hx, hy)` of fd2d_3_4.py:151-161: the D
 * update (as fd2d_dfield) and the incident source ezi[inc_pos] = inc_val
 * (the reference's ezi[3] = gaussian(t, 20, 8)).                            */
int fd2d_dfield_inc(int t, int nx, int ny, const fdtd_pml *pml, float *ezi, float *dz,
                    const float *hx, const float *hy, int inc_pos, float inc_val, void *stream);
/* `inctdz(nx, ny, npml, hxi, dz)` (fd2d_3_4.py:164-170).                   */
int fd2d_inctdz(int nx, int ny, int npml, const float *hxi, float *dz, void *stream);
/* `hxinct(ny, ezi, hxi)` (fd2d_3_4.py:186-191).                             */
int fd2d_hxinct(int ny, const float *ezi, float *hxi, void *stream);
/* `incthx(nx, ny, npml, ezi, hx)` and `incthy(nx, ny, npml, ezi, hy)`
 * (fd2d_3_4.py:211-226).                                                    */
int fd2d_incthx(int nx, int ny, int npml, const float *ezi, float *hx, void *stream);
int fd2d_incthy(int nx, int ny, int npml, const float *ezi, float *hy, void *stream);
/* `fourier(t, nf, nx, ny, dt, freq, ezi, ez, ft)` (fd2d_3_4.py:121-136):
 * freq is a host array of nf <= FDTD_MAX_FREQ frequencies; the twiddles
 * cos/sin(2*pi*freq*dt*t) are evaluated on the host (double product,
 * narrowed to float, cosf/sinf) as fd2d/clang/test_3_4.c:42-49 evaluates
 * them; the incident probe is ezi[6], as in the reference.                  */
int fd2d_fourier(int t, int nf, int nx, int ny, float dt, const float *freq, const float *ezi,
                 const float *ez, const fdtd_ftrans *ft, void *stream);

/* A whole plane-wave run: per step t = t0+1 .. t0+nsteps the sequence of
 * fd2d/clang/test_3_4.c:223-233 (ezinct with ezi[inc_pos] = inc_wave[t-1];
 * D update with d's PML and hard sources; inctdz; E update with d->md;
 * fourier if nf > 0; hxinct; H update; incthx; incthy; d's probes).  Uses
 * the caller's state set of d (no partner set needed).                     */
typedef struct fd2d_tfsf_desc {
    int npml;                   /* box rows/columns npml-1 .. n-npml (>= 2)  */
    float *ezi, *hxi, *bc;      /* incident grid, ny cells each; bc[4]       */
    int inc_pos;                /* incident source cell (3 in the reference) */
    const float *inc_wave;      /* host: inc_wave[t-1] for step t            */
    int inc_len;
    int nf;                     /* DFT frequencies (0: no DFT)               */
    const float *freq;          /* host, nf values                           */
    float dt;
    fdtd_ftrans ft;
    int dft_probe;              /* incident probe cell (6 in the reference)  */
    float *work;                /* NULL: the per-step kernels; else fused     */
    size_t work_elems;          /* >= fd2d_tfsf_work_elems(ny, npml)          */
} fd2d_tfsf_desc;
/* With work, the run is fused: per chunk of up to 64 steps the incident grid
 * advances in one workgroup and records what the box needs per step into
 * work; the temporally blocked launches (d's partner set and launch shape,
 * as fd2d_tm_run, at most 8 steps per launch) apply the TF/SF corrections
 * and accumulate the DFT (PML form, lossy or not, nf <= 3), with the same
 * bits as the per-step path.                                                */
size_t fd2d_tfsf_work_elems(int ny, int npml);
int fd2d_tfsf_run(const fd2d_tm_desc *d, const fd2d_tfsf_desc *b, int t0, int nsteps,
                  void *stream);

/* Ridge least squares for the linear path.  lbc_gram accumulates, over the
 * rows (as lbc_train_grad: targets tgt[r][3], or from the next sample when
 * tgt is NULL) with s > 0, the FP64
 * sums G = sum w xn xn^T (108 x 108) and R = sum w xn t^T (108 x 3) of the
 * normalised inputs xn (bf16-rounded when round_bf16) and residual targets
 * t = (T - base) / s with row weights w = s^wpow (0 <= wpow <= 4; 0: all
 * rows alike, 2: least squares on the raw values), plus the weight sum
 * (the count for wpow = 0): nblocks workgroups write
 * part[block][lbc_gram_elems()] (row-major [G | R], then the count).
 * lbc_fit_linear reduces the partials in a fixed order into gram, then (if
 * params is not NULL) solves (G + rho * wsum * I) L^T = R by Cholesky in
 * FP64 and writes L into params; status[0] = 0, or 1 + k if pivot k is not
 * positive (L then unchanged).  Call lbc_pack afterwards only if the bf16
 * buffer must follow (L is read in FP32).                                   */
int lbc_gram_elems(void);
int lbc_gram(int n, const float *src, const long long *offs, long long tstride, const float *tgt,
             int round_bf16, float wpow, int nblocks, double *part, void *stream);
int lbc_fit_linear(int nblocks, const double *part, double *gram, double rho, float *params,
                   int *status, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FDTD_H */
