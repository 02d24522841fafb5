// fdtd_capi.hip -- the extern "C" entry points of libfdtd_hip.so (include/fdtd.h).
//
// Each entry point validates its arguments on the host before queueing any
// device work (FDTD_EINVAL + message on rejection), then launches the kernels
// of fd1d.hip / fd2d.hip / fd2d_tb.hip on the caller's stream.  Nothing here
// allocates device memory.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <climits>
#include <mutex>
#include <vector>

#include "fdtd_internal.h"
#include "lbc_impl.h"

namespace fdtd {

static thread_local char g_err[512];

void clear_error() { g_err[0] = '\0'; }

int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return FDTD_EINVAL;
}

int hip_check(hipError_t e, const char *what)
{
    if (e == hipSuccess) return 0;
    snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
    return (int)e;
}

namespace {

inline hipStream_t S(void *s) { return reinterpret_cast<hipStream_t>(s); }

bool pml_all(const fdtd_pml *p)
{
    return p && p->fx1 && p->fx2 && p->fx3 && p->fy1 && p->fy2 && p->fy3 && p->gx2 && p->gx3 &&
           p->gy2 && p->gy3;
}

bool pml_none(const fdtd_pml *p)
{
    return !p || (!p->fx1 && !p->fx2 && !p->fx3 && !p->fy1 && !p->fy2 && !p->fy3 && !p->gx2 &&
                  !p->gx3 && !p->gy2 && !p->gy3);
}

PmlPtrs to_ptrs(const fdtd_pml &p)
{
    return PmlPtrs{p.fx1, p.fx2, p.fx3, p.fy1, p.fy2, p.fy3, p.gx2, p.gx3, p.gy2, p.gy3};
}

template <typename T>
int check_fd1d_ex(int nx, const T *ca, const T *cb, T *ex, const T *hy, int mode, int pos)
{
    if (nx < 2) return fail("fd1d_exfield: nx=%d < 2", nx);
    if (!ex || !hy) return fail("fd1d_exfield: ex and hy are required");
    if (ca && !cb) return fail("fd1d_exfield: ca given without cb");
    if (mode < FDTD_SRC_NONE || mode > FDTD_SRC_ADD) return fail("fd1d_exfield: bad src_mode %d", mode);
    if (mode != FDTD_SRC_NONE && (pos < 0 || pos >= nx))
        return fail("fd1d_exfield: src_pos %d outside [0,%d)", pos, nx);
    return 0;
}

template <typename T>
int fd1d_run_impl(int nx, int t0, int nsteps, const T *ca, const T *cb, T *ex, T *hy, T *bc,
                  int mode, int pos, const T *wave, int nprobe, const int *probe_k,
                  T *probe_out, void *stream)
{
    clear_error();
    int rc = check_fd1d_ex(nx, ca, cb, ex, hy, mode, pos);
    if (rc) return rc;
    if (nx < 3) return fail("fd1d_run: nx=%d < 3", nx);
    if (t0 < 0 || nsteps < 0) return fail("fd1d_run: t0=%d nsteps=%d", t0, nsteps);
    if (mode != FDTD_SRC_NONE && !wave) return fail("fd1d_run: src_wave required for a source");
    if (nprobe < 0 || nprobe > FDTD_MAX_PROBES) return fail("fd1d_run: nprobe=%d", nprobe);
    if (nprobe && (!probe_k || !probe_out)) return fail("fd1d_run: probe_k/probe_out required");
    for (int p = 0; p < nprobe; ++p)
        if (probe_k[p] < 0 || probe_k[p] >= nx) return fail("fd1d_run: probe %d at %d", p, probe_k[p]);
    if ((long)2 * sizeof(T) * nx > fd1d_run_lds_limit())
        return fail("fd1d_run: nx=%d exceeds the single-workgroup limit %d", nx,
                    fd1d_run_max_nx((int)sizeof(T)));
    if (nsteps == 0) return 0;
    return launch_fd1d_run<T>(nx, nsteps, ca, cb, ex, hy, bc, mode, pos, wave, nprobe, probe_k,
                              probe_out, S(stream));
}

// ---------------------------------------------------------- 2-D TM plan

struct Plan {
    int depth, rows, cols, halo;
    int launches;          // ping-pong launches
    int tail_unfused;      // 1 if the last step runs as dfield/efield/hfield in place
    int mw;                // 1: multi-wave form (rows = rows of its interior blocks)
    int frows;             // rows of the single-wave tiles (strips, or the frame of the
                           // multi-wave form)
};

// Rows of an interior block of the multi-wave form when the caller gives
// none.
int mw_default_rows(int nx)
{
    long h = nx / 32;
    if (h < 32) h = 32;
    if (h > 512) h = 512;
    return (int)h;
}

// Default depth and rows of the single-wave strips, from sweeps on MI355X
// (lossy PML form, one column per lane; Mcell-updates/s at 1000-2000 steps):
//   2048^2:  depth 10, rows 12 (355k; rows 16/24: 260k/206k; depth 8 rows 16:
//            278k; depth 11 rows 12/16: 342k/305k)
//   4096^2:  depth 11, rows 32 (621k; rows 40: 626k; depth 10 rows 32/42:
//            592k/595k; depth 8 rows 16/32/48: 449k/494k/461k)
//   8192^2:  depth 11, rows 64 (737k-757k; rows 56/72/85/96: 699k/732k/724k/
//            725k; depth 10 rows 64: 698k-715k; depth 8 rows 64: 586k;
//            depth 12: 576k)
//   16384^2: depth 11, rows 96 (815k; rows 128: 817k; depth 10 rows 96:
//            774k; depth 8 rows 96: 606k)
// Depths 10 and 11 keep two carry sets in 3 waves per SIMD (depth 12 needs
// 192 VGPRs: 2 waves); depth 11 moves the arrays once per 11 steps instead
// of 10 and wins from 4096^2 on, where its one more halo column and warm-up
// row per side cost less.  Rows follow nx/128 for 4096 <= nx < 12288 and
// ~nx/170 outside.  With the free-space tiles that move no iz (fewer bytes
// per launch) the two depths are close below 16384^2: 8192^2 depth 10/11
// rows 64: 852k/839k, rows 85: 831k/841k; 4096^2 rows 32: 679k/684k;
// 16384^2 rows 96: 855k/886k (the autotuner picks per grid).
void tm_defaults(int nx, int *depth, int *rows)
{
    *depth = nx >= 4096 ? 11 : (nx >= 2048 ? 10 : 8);
    long h = nx >= 4096 && nx < 12288 ? nx / 128 : nx / 170;
    if (h < 8) h = 8;
    if (h > 512) h = 512;
    *rows = (int)h;
}

int plan_tm(const fd2d_tm_desc *d, int nsteps, Plan *pl)
{
    int D0 = 8, H0 = 16;
    tm_defaults(d->nx, &D0, &H0);
    int D = d->depth ? d->depth : (d->form == FDTD_TM_FORM_MULTIWAVE ? 8 : D0);
    int C = d->cols ? d->cols : 1;
    if (!tb_supported(D, C)) return fail("fd2d_tm: depth=%d cols=%d not supported", D, C);
    if (64 * C - 2 * D < 8) return fail("fd2d_tm: depth %d too large for %d columns/lane", D, C);
    if (d->form != FDTD_TM_FORM_AUTO && d->form != FDTD_TM_FORM_STRIP &&
        d->form != FDTD_TM_FORM_MULTIWAVE)
        return fail("fd2d_tm: form=%d", d->form);
    if (d->form == FDTD_TM_FORM_MULTIWAVE && (C != 1 || D > MW_MAX_DEPTH))
        return fail("fd2d_tm: the multi-wave form takes one column per lane and depth <= %d "
                    "(cols=%d depth=%d)", MW_MAX_DEPTH, C, D);
    const long h = H0;
    const int mw = d->form == FDTD_TM_FORM_MULTIWAVE;
    int H = d->rows ? d->rows : (mw ? mw_default_rows(d->nx) : (int)h);
    if (H < 1) return fail("fd2d_tm: rows=%d", H);
    // strip halo: the deepest launch's cone, rounded up to whole lanes (set
    // again below from the launches of this call)
    pl->depth = D; pl->cols = C; pl->rows = H; pl->halo = ((D + C - 1) / C) * C;
    pl->mw = mw; pl->frows = mw ? (int)h : H;
    pl->launches = 0; pl->tail_unfused = 0;
    if (nsteps <= 0) return 0;
    if (nsteps == 1) { pl->tail_unfused = 1; return 0; }
    int L = (nsteps + D - 1) / D;
    if (L % 2) {
        if (L + 1 <= nsteps) L += 1;
        else { pl->tail_unfused = 1; L = (nsteps - 1 + D - 1) / D; if (L % 2) L += 1; }
    }
    pl->launches = L;
    // the strip halo of this call: its deepest launch's cone (a 20-step call
    // at depth 11 runs two launches of 10 steps, which need 10 halo columns,
    // not 11)
    const int nfused = nsteps - pl->tail_unfused;
    const int kmax = nfused / L + (nfused % L ? 1 : 0);
    pl->halo = ((kmax + C - 1) / C) * C;
    return 0;
}

// The learned boundary's two extra streams (the boundary part of the blocked
// TM launch and the frame kernel run concurrently with the interior part on
// the caller's stream) and the events that order them.  One set per host
// thread and device, created on first use and kept for the life of the
// process (no device memory is involved).
struct Aux {
    hipStream_t s = nullptr, s2 = nullptr;
    hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr, join0 = nullptr;
};

int aux_for_current_device(Aux **out)
{
    static thread_local Aux tab[64];
    int dev = 0;
    int rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (rc) return rc;
    if (dev < 0 || dev >= 64) return fail("device ordinal %d out of range", dev);
    Aux &a = tab[dev];
    if (!a.s) {
        rc = hip_check(hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking), "aux stream");
        if (!rc) rc = hip_check(hipStreamCreateWithFlags(&a.s2, hipStreamNonBlocking), "aux stream");
        // events order work between streams of this device only: a
        // device-scope release is enough (no system-scope fence)
        const unsigned ef = hipEventDisableTiming | hipEventReleaseToDevice;
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.fork, ef), "aux event");
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.join, ef), "aux event");
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.join2, ef), "aux event");
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.join0, ef), "aux event");
        if (rc) return rc;
    }
    *out = &a;
    return 0;
}

// ------------------------------------------------------- batched grids

struct BatchGeom {
    int depth, rows, cols, halo;   // rows 0: the default (batch_grid_rows)
};

int batch_geom(int depth, int rows, int cols, BatchGeom *g)
{
    g->depth = depth ? depth : 8;
    g->cols = cols ? cols : 1;
    g->rows = rows;
    if (!tb_supported(g->depth, g->cols))
        return fail("fd2d_batch: depth=%d cols=%d not supported", g->depth, g->cols);
    g->halo = ((g->depth + g->cols - 1) / g->cols) * g->cols;
    if (64 * g->cols - 2 * g->halo < 8) return fail("fd2d_batch: depth too large for the strip");
    if (g->rows < 0) return fail("fd2d_batch: rows=%d", g->rows);
    return 0;
}

// Rows of a grid's blocks: the caller's, or (rows = 0) 128 -- on MI355X the
// 256-grid batch of sides 128-1024, with the units dealt out as below, ran
// 374k Mcell-updates/s at 128 rows, 360k at 112, 343k at 160, 343k at 64
// (and, dealt out in list order, 330k at 64 and 329k at 128).
int batch_grid_rows(const BatchGeom &g, int nx)
{
    const int h = g.rows ? g.rows : 128;
    return h < nx ? h : nx;
}

int check_grid(const fd2d_grid &G, int k)
{
    if (G.nx < 2 || G.ny < 2) return fail("fd2d_batch: grid %d is %dx%d", k, G.nx, G.ny);
    if ((long long)G.nx * G.ny >= (1LL << 30)) return fail("fd2d_batch: grid %d too large", k);
    if (!G.dz || !G.iz || !G.ez || !G.hx || !G.hy || !G.ihx || !G.ihy || !G.dz2 || !G.iz2 ||
        !G.hx2 || !G.hy2 || !G.ihx2 || !G.ihy2)
        return fail("fd2d_batch: grid %d: both state sets are required", k);
    if (!G.md.naz || !G.md.nbz || !pml_all(&G.pml))
        return fail("fd2d_batch: grid %d: the lossy PML form (naz, nbz, ten profiles) is required", k);
    if (G.src_i >= 0 && (G.src_i >= G.nx || G.src_j < 0 || G.src_j >= G.ny || !G.src_wave))
        return fail("fd2d_batch: grid %d: source outside the grid or without a waveform", k);
    if (G.probe_i >= 0 && (G.probe_i >= G.nx || G.probe_j < 0 || G.probe_j >= G.ny || !G.probe_out))
        return fail("fd2d_batch: grid %d: probe outside the grid or without output", k);
    return 0;
}

// launches of one batched run (shared by fd2d_batch_run and the graph capture)
int batch_launches(const fd2d_batch_desc *b, int t0, int nsteps, int max_wave_len,
                   int max_probe_len, hipStream_t s)
{
    if (!b || !b->grids_dev || !b->units_dev || b->nunits <= 0 || b->waves <= 0)
        return fail("fd2d_batch_run: incomplete batch descriptor");
    if (!tb_supported(b->depth, b->cols)) return fail("fd2d_batch_run: depth/cols");
    if (t0 < 0 || nsteps < 0) return fail("fd2d_batch_run: t0=%d nsteps=%d", t0, nsteps);
    // the device indexes each grid's waveform and probe trace at t-1 with
    // that grid's own lengths: the recorded minima bound every grid
    if (t0 + nsteps > b->wave_len || t0 + nsteps > b->probe_len)
        return fail("fd2d_batch_run: steps up to %d exceed a grid's waveform (%d) or probe (%d) length",
                    t0 + nsteps, b->wave_len, b->probe_len);
    if (t0 + nsteps > max_wave_len || t0 + nsteps > max_probe_len)
        return fail("fd2d_batch_run: steps up to %d exceed a waveform or probe length", t0 + nsteps);
    if (nsteps == 0) return 0;
    const int D = b->depth;
    int L = (nsteps + D - 1) / D;
    if (L % 2) L += 1;
    if (L > nsteps)
        return fail("fd2d_batch_run: %d steps cannot be split into an even number of launches", nsteps);
    const int halo = ((D + b->cols - 1) / b->cols) * b->cols;
    // free-space units: the first launch (the deepest: its units read every
    // row a later launch's units read) records every unit's entry with a tag
    // fresh to the call (bit 0 clear), the later launches use them
    BatchVac vac{b->unit_work, 0, 0u};
    if (b->unit_work) {
        static std::atomic<unsigned> calls{0};
        vac.tag = (calls.fetch_add(1) + 1u) << 1;
    }
    int done = 0;
    for (int l = 0; l < L; ++l) {
        const int k = nsteps / L + (l < nsteps % L ? 1 : 0);
        vac.mode = b->unit_work ? (l == 0 ? 1 : 2) : 0;
        int rc = launch_fd2d_batch(b->grids_dev, b->units_dev, b->nunits, b->waves, k, b->cols,
                                   b->vec != 0, l % 2, t0 + done + 1, l == L - 1, halo, vac, s);
        if (rc) return rc;
        done += k;
    }
    return 0;
}

}  // namespace
}  // namespace fdtd

using namespace fdtd;

extern "C" {

int fdtd_abi_version(void) { return FDTD_ABI_VERSION; }

const char *fdtd_last_error(void) { return g_err; }

int fdtd_preload(void)
{
    clear_error();
    int rc = preload_fd1d();
    if (!rc) rc = preload_fd2d();
    if (!rc) rc = preload_tb();
    if (!rc) rc = preload_mw();
    if (!rc) rc = preload_dform1d();
    if (!rc) rc = preload_tfsf();
    if (!rc) rc = preload_lbc();
    if (!rc) rc = preload_lbc_frame();
    double *scr = nullptr;          // device address of the energy scratch
    if (!rc) rc = energy_scratch(&scr);
    return rc;
}

// ------------------------------------------------------------ waveforms

float fdtd_gaussian_f32(int t, int t0, float sigma)
{
    return expf(-0.5f * (t - t0) / sigma * (t - t0) / sigma);
}

double fdtd_gaussian_f64(int t, int t0, double sigma)
{
    return exp(-0.5 * (t - t0) / sigma * (t - t0) / sigma);
}

float fdtd_sinusoidal_f32(int t, float ds, float freq)
{
    float dt = ds / 6e8f;
    return sinf(2 * M_PI * freq * dt * t);
}

double fdtd_sinusoidal_f64(int t, double ds, double freq)
{
    double dt = ds / 6e8;
    return sin(2 * M_PI * freq * dt * t);
}

// ------------------------------------------------------------------ 1-D

int fd1d_exfield(int t, int nx, const float *ca, const float *cb, float *ex, const float *hy,
                 int src_mode, int src_pos, float src_val, void *stream)
{
    (void)t;
    clear_error();
    int rc = check_fd1d_ex(nx, ca, cb, ex, hy, src_mode, src_pos);
    if (rc) return rc;
    return launch_fd1d_exfield<float>(nx, ca, cb, ex, hy, src_mode, src_pos, src_val, S(stream));
}

int fd1d_exfield_f64(int t, int nx, const double *ca, const double *cb, double *ex,
                     const double *hy, int src_mode, int src_pos, double src_val, void *stream)
{
    (void)t;
    clear_error();
    int rc = check_fd1d_ex(nx, ca, cb, ex, hy, src_mode, src_pos);
    if (rc) return rc;
    return launch_fd1d_exfield<double>(nx, ca, cb, ex, hy, src_mode, src_pos, src_val, S(stream));
}

int fd1d_hyfield(int nx, float *ex, float *hy, float *bc, void *stream)
{
    clear_error();
    if (nx < 3) return fail("fd1d_hyfield: nx=%d < 3", nx);
    if (!ex || !hy) return fail("fd1d_hyfield: ex and hy are required");
    return launch_fd1d_hyfield<float>(nx, ex, hy, bc, S(stream));
}

int fd1d_hyfield_f64(int nx, double *ex, double *hy, double *bc, void *stream)
{
    clear_error();
    if (nx < 3) return fail("fd1d_hyfield: nx=%d < 3", nx);
    if (!ex || !hy) return fail("fd1d_hyfield: ex and hy are required");
    return launch_fd1d_hyfield<double>(nx, ex, hy, bc, S(stream));
}

int fd1d_run(int nx, int t0, int nsteps, const float *ca, const float *cb, float *ex, float *hy,
             float *bc, int src_mode, int src_pos, const float *src_wave, int nprobe,
             const int *probe_k, float *probe_out, void *stream)
{
    return fd1d_run_impl<float>(nx, t0, nsteps, ca, cb, ex, hy, bc, src_mode, src_pos, src_wave,
                                nprobe, probe_k, probe_out, stream);
}

int fd1d_run_f64(int nx, int t0, int nsteps, const double *ca, const double *cb, double *ex,
                 double *hy, double *bc, int src_mode, int src_pos, const double *src_wave,
                 int nprobe, const int *probe_k, double *probe_out, void *stream)
{
    return fd1d_run_impl<double>(nx, t0, nsteps, ca, cb, ex, hy, bc, src_mode, src_pos, src_wave,
                                 nprobe, probe_k, probe_out, stream);
}

int fd1d_run_max_nx(int elem_size)
{
    if (elem_size <= 0) return 0;
    return fd1d_run_lds_limit() / (2 * elem_size);
}

int fd1d_dielectric(int nx, double dt, double epsr, double sigma, float *ca, float *cb,
                    void *stream)
{
    clear_error();
    if (nx < 1 || !ca || !cb) return fail("fd1d_dielectric: nx=%d, ca/cb required", nx);
    if (!(epsr > 0)) return fail("fd1d_dielectric: epsr must be > 0");
    return launch_fd1d_dielectric<float>(nx, dt, epsr, sigma, ca, cb, S(stream));
}

int fd1d_dielectric_f64(int nx, double dt, double epsr, double sigma, double *ca, double *cb,
                        void *stream)
{
    clear_error();
    if (nx < 1 || !ca || !cb) return fail("fd1d_dielectric: nx=%d, ca/cb required", nx);
    if (!(epsr > 0)) return fail("fd1d_dielectric: epsr must be > 0");
    return launch_fd1d_dielectric<double>(nx, dt, epsr, sigma, ca, cb, S(stream));
}

// ------------------------------------------------------------------ 2-D

int fd2d_dfield(int t, int nx, int ny, const fdtd_pml *pml, float *dz, const float *hx,
                const float *hy, int nsrc, const int *src_i, const int *src_j,
                const float *src_val, void *stream)
{
    (void)t;
    clear_error();
    if (nx < 2 || ny < 2) return fail("fd2d_dfield: grid %dx%d too small", nx, ny);
    if (!dz || !hx || !hy) return fail("fd2d_dfield: dz, hx, hy are required");
    if (!pml_none(pml) && !pml_all(pml)) return fail("fd2d_dfield: pml must be complete or absent");
    if (nsrc < 0 || nsrc > FDTD_MAX_SOURCES) return fail("fd2d_dfield: nsrc=%d", nsrc);
    if (nsrc && (!src_i || !src_j || !src_val)) return fail("fd2d_dfield: source arrays required");
    SrcList src{};
    src.n = nsrc;
    for (int k = 0; k < nsrc; ++k) {
        if (src_i[k] < 0 || src_i[k] >= nx || src_j[k] < 0 || src_j[k] >= ny)
            return fail("fd2d_dfield: source %d at (%d,%d) outside the grid", k, src_i[k], src_j[k]);
        src.i[k] = src_i[k]; src.j[k] = src_j[k]; src.v[k] = src_val[k];
    }
    PmlPtrs pp;
    const PmlPtrs *ppp = nullptr;
    if (pml_all(pml)) { pp = to_ptrs(*pml); ppp = &pp; }
    return launch_fd2d_dfield(nx, ny, ppp, dz, hx, hy, src, S(stream));
}

int fd2d_efield(int nx, int ny, const fdtd_medium *md, const float *dz, float *iz, float *ez,
                void *stream)
{
    clear_error();
    if (nx < 1 || ny < 1) return fail("fd2d_efield: grid %dx%d", nx, ny);
    if (!md || !md->naz) return fail("fd2d_efield: md->naz is required");
    if (!dz || !ez) return fail("fd2d_efield: dz and ez are required");
    if (md->nbz && !iz) return fail("fd2d_efield: iz is required with nbz");
    return launch_fd2d_efield(nx, ny, md->naz, md->nbz, dz, iz, ez, S(stream));
}

int fd2d_hfield(int nx, int ny, const fdtd_pml *pml, const float *ez, float *ihx, float *ihy,
                float *hx, float *hy, void *stream)
{
    clear_error();
    if (nx < 2 || ny < 2) return fail("fd2d_hfield: grid %dx%d too small", nx, ny);
    if (!ez || !hx || !hy) return fail("fd2d_hfield: ez, hx, hy are required");
    if (!pml_none(pml) && !pml_all(pml)) return fail("fd2d_hfield: pml must be complete or absent");
    PmlPtrs pp;
    const PmlPtrs *ppp = nullptr;
    if (pml_all(pml)) {
        if (!ihx || !ihy) return fail("fd2d_hfield: ihx, ihy are required with a pml");
        pp = to_ptrs(*pml);
        ppp = &pp;
    }
    return launch_fd2d_hfield(nx, ny, ppp, ez, ihx, ihy, hx, hy, S(stream));
}

int fd2d_pmlparam(int nx, int ny, int npml, const fdtd_pml *pml, void *stream)
{
    clear_error();
    if (!pml_all(pml)) return fail("fd2d_pmlparam: all ten profile arrays are required");
    if (nx < 2 || ny < 2) return fail("fd2d_pmlparam: grid %dx%d too small", nx, ny);
    const int m = nx < ny ? nx : ny;
    if (npml < 0 || npml > m - 2) return fail("fd2d_pmlparam: npml=%d outside [0,%d]", npml, m - 2);
    return launch_fd2d_pmlparam(nx, ny, npml, *pml, S(stream));
}

int fd2d_dielectric(int nx, int ny, int npml, double dt, int ncyl, const double *cyl,
                    int int_sub, float *naz, float *nbz, void *stream)
{
    clear_error();
    if (nx < 1 || ny < 1 || !naz || !nbz) return fail("fd2d_dielectric: grid/naz/nbz");
    if (npml < 0) return fail("fd2d_dielectric: npml=%d", npml);
    if (ncyl < 0 || ncyl > FDTD_MAX_CYLINDERS) return fail("fd2d_dielectric: ncyl=%d", ncyl);
    if (ncyl && !cyl) return fail("fd2d_dielectric: cyl required");
    CylList cl{};
    cl.n = ncyl;
    for (int c = 0; c < ncyl; ++c) {
        cl.cx[c] = cyl[5 * c + 0]; cl.cy[c] = cyl[5 * c + 1]; cl.rad[c] = cyl[5 * c + 2];
        cl.epsr[c] = cyl[5 * c + 3]; cl.sigma[c] = cyl[5 * c + 4];
        if (!(cl.rad[c] >= 0)) return fail("fd2d_dielectric: cylinder %d radius", c);
    }
    return launch_fd2d_dielectric(nx, ny, npml, dt, cl, int_sub, naz, nbz, S(stream));
}

int fd2d_energy(int nx, int ny, int i0, int i1, int j0, int j1, const float *ez,
                const float *hx, const float *hy, double *work, double *result, void *stream)
{
    clear_error();
    if (!ez || !hx || !hy || !work) return fail("fd2d_energy: ez, hx, hy, work required");
    if (i0 < 0 || j0 < 0 || i1 > nx || j1 > ny || i0 > i1 || j0 > j1)
        return fail("fd2d_energy: window [%d,%d)x[%d,%d) outside %dx%d", i0, i1, j0, j1, nx, ny);
    int rc = launch_fd2d_energy(nx, ny, i0, i1, j0, j1, ez, hx, hy, work, S(stream));
    if (rc || !result) return rc;
    rc = hip_check(hipMemcpyAsync(result, work, sizeof(double), hipMemcpyDeviceToHost, S(stream)),
                   "fd2d_energy copy");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize(S(stream)), "fd2d_energy sync");
}

int fd2d_energy_noalloc(int nx, int ny, int i0, int i1, int j0, int j1, const float *ez,
                        const float *hx, const float *hy, double *result, void *stream)
{
    clear_error();
    if (!ez || !hx || !hy || !result) return fail("fd2d_energy_noalloc: ez, hx, hy, result required");
    if (nx < 1 || ny < 1) return fail("fd2d_energy_noalloc: grid %dx%d", nx, ny);
    if (i0 < 0 || j0 < 0 || i1 > nx || j1 > ny || i0 > i1 || j0 > j1)
        return fail("fd2d_energy_noalloc: window [%d,%d)x[%d,%d) outside %dx%d", i0, i1, j0, j1, nx, ny);
    // the partial sums need 1 + 1024 doubles whatever the window: the
    // scratch holds exactly that, so no window can exceed it
    double *scr = nullptr;
    int rc = energy_scratch(&scr);
    if (rc) return rc;
    return launch_fd2d_energy_to(nx, ny, i0, i1, j0, j1, ez, hx, hy, scr, result, S(stream));
}

int fd2d_energy_diff(int wn, int wm, int nxa, int nya, int ia0, int ja0, const float *eza,
                     const float *hxa, const float *hya, int nxb, int nyb, int ib0, int jb0,
                     const float *ezb, const float *hxb, const float *hyb, double *work,
                     double *result, void *stream)
{
    clear_error();
    if (!eza || !hxa || !hya || !ezb || !hxb || !hyb || !work)
        return fail("fd2d_energy_diff: null pointer");
    if (wn < 0 || wm < 0 || ia0 < 0 || ja0 < 0 || ib0 < 0 || jb0 < 0 || ia0 + wn > nxa ||
        ja0 + wm > nya || ib0 + wn > nxb || jb0 + wm > nyb)
        return fail("fd2d_energy_diff: window %dx%d at (%d,%d)/(%d,%d) outside %dx%d / %dx%d", wn,
                    wm, ia0, ja0, ib0, jb0, nxa, nya, nxb, nyb);
    int rc = launch_fd2d_energy_diff(wn, wm, nya, ia0, ja0, eza, hxa, hya, nyb, ib0, jb0, ezb, hxb,
                                     hyb, work, S(stream));
    if (rc || !result) return rc;
    rc = hip_check(hipMemcpyAsync(result, work, sizeof(double), hipMemcpyDeviceToHost, S(stream)),
                   "fd2d_energy_diff copy");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize(S(stream)), "fd2d_energy_diff sync");
}

static int tm_validate(const fd2d_tm_desc *d, int t0, int nsteps)
{
    if (!d) return fail("fd2d_tm: desc is NULL");
    if (d->nx < 2 || d->ny < 2) return fail("fd2d_tm: grid %dx%d too small", d->nx, d->ny);
    if ((long long)d->nx * d->ny >= (1LL << 30))
        return fail("fd2d_tm: grid %dx%d has 2^30 cells or more (32-bit offsets)", d->nx, d->ny);
    if (!d->dz || !d->ez || !d->hx || !d->hy || !d->md.naz)
        return fail("fd2d_tm: dz, ez, hx, hy, md.naz are required");
    const bool lossy = d->md.nbz != nullptr;
    const bool pml = pml_all(&d->pml);
    if (!pml && !pml_none(&d->pml)) return fail("fd2d_tm: pml must be complete or absent");
    if (lossy && !d->iz) return fail("fd2d_tm: iz is required with nbz");
    if (pml && (!d->ihx || !d->ihy)) return fail("fd2d_tm: ihx, ihy are required with a pml");
    if (t0 < 0 || nsteps < 0) return fail("fd2d_tm: t0=%d nsteps=%d", t0, nsteps);
    if (d->nsrc < 0 || d->nsrc > FDTD_MAX_SOURCES) return fail("fd2d_tm: nsrc=%d", d->nsrc);
    if (d->nsrc) {
        if (!d->src_wave) return fail("fd2d_tm: src_wave required");
        if (t0 + nsteps > d->wave_len)
            return fail("fd2d_tm: steps up to %d exceed wave_len %d", t0 + nsteps, d->wave_len);
        for (int k = 0; k < d->nsrc; ++k)
            if (d->src_i[k] < 0 || d->src_i[k] >= d->nx || d->src_j[k] < 0 || d->src_j[k] >= d->ny)
                return fail("fd2d_tm: source %d outside the grid", k);
    }
    if (d->nprobe < 0 || d->nprobe > FDTD_MAX_PROBES) return fail("fd2d_tm: nprobe=%d", d->nprobe);
    if (d->nprobe) {
        if (!d->probe_out) return fail("fd2d_tm: probe_out required");
        if (t0 + nsteps > d->probe_len)
            return fail("fd2d_tm: steps up to %d exceed probe_len %d", t0 + nsteps, d->probe_len);
        for (int p = 0; p < d->nprobe; ++p)
            if (d->probe_i[p] < 0 || d->probe_i[p] >= d->nx || d->probe_j[p] < 0 ||
                d->probe_j[p] >= d->ny)
                return fail("fd2d_tm: probe %d outside the grid", p);
    }
    return 0;
}

int fd2d_tm_plan(const fd2d_tm_desc *d, int nsteps, int *launches, int *depth, int *rows,
                 int *cols)
{
    clear_error();
    int rc = tm_validate(d, 0, 0);
    if (rc) return rc;
    Plan pl;
    rc = plan_tm(d, nsteps, &pl);
    if (rc) return rc;
    if (launches) *launches = pl.launches + (pl.tail_unfused ? 1 : 0);
    if (depth) *depth = pl.depth;
    if (rows) *rows = pl.rows;
    if (cols) *cols = pl.cols;
    return 0;
}

// Free-space tile flags recorded per workspace (host side): the tag they were
// written with and what they describe.
struct TileRecord {
    const unsigned *work;
    const float *naz, *nbz, *iz, *iz2;
    int nx, ny, rows, cols, halo, depth;
    unsigned tag;
    bool same_tiles(const TileRecord &o) const
    {
        return naz == o.naz && nbz == o.nbz && iz == o.iz && iz2 == o.iz2 && nx == o.nx &&
               ny == o.ny && rows == o.rows && cols == o.cols && halo == o.halo;
    }
};

static std::mutex &tile_records_mutex()
{
    static std::mutex m;
    return m;
}

static std::vector<TileRecord> &tile_records()
{
    static std::vector<TileRecord> v;
    return v;
}

// Tiles of the single-wave strip form of a plan: strips x row blocks.
static size_t tm_tiles(const fd2d_tm_desc *d, const Plan &pl)
{
    if (pl.mw) return 0;
    const int w = 64 * pl.cols - 2 * pl.halo;
    return (size_t)((d->ny + w - 1) / w) * (size_t)((d->nx + pl.rows - 1) / pl.rows);
}

size_t fd2d_tm_tile_elems(const fd2d_tm_desc *d, int nsteps)
{
    clear_error();
    if (tm_validate(d, 0, 0)) return 0;
    Plan pl;
    if (plan_tm(d, nsteps, &pl)) return 0;
    return tm_tiles(d, pl);
}

int fd2d_tm_run(const fd2d_tm_desc *d, int t0, int nsteps, void *stream)
{
    clear_error();
    int rc = tm_validate(d, t0, nsteps);
    if (rc) return rc;
    Plan pl;
    rc = plan_tm(d, nsteps, &pl);
    if (rc) return rc;
    if (nsteps == 0) return 0;
    const bool lossy = d->md.nbz != nullptr;
    const bool pml = pml_all(&d->pml);
    if (pl.launches) {
        if (!d->dz2 || !d->hx2 || !d->hy2 || (lossy && !d->iz2) || (pml && (!d->ihx2 || !d->ihy2)))
            return fail("fd2d_tm: the partner state set (dz2, iz2, hx2, hy2, ihx2, ihy2) is required");
    }
    const int nfused = nsteps - pl.tail_unfused;
    TbParams p{};
    p.nx = d->nx; p.ny = d->ny;
    p.naz = d->md.naz; p.nbz = d->md.nbz;
    if (pml) p.pml = to_ptrs(d->pml);
    p.halo = pl.halo; p.rows = pl.rows;
    // interior band: inside [1, n-2] and, with a PML, inside the promised
    // default-profile band (empty promise: every wave runs the general body)
    if (pml) {
        const int *br = d->band_rows, *bc = d->band_cols;
        p.rlo = br[1] > br[0] ? (br[0] > 1 ? br[0] : 1) : 0;
        p.rhi = br[1] > br[0] ? (br[1] < d->nx - 1 ? br[1] : d->nx - 1) : 0;
        p.clo = bc[1] > bc[0] ? (bc[0] > 1 ? bc[0] : 1) : 0;
        p.chi = bc[1] > bc[0] ? (bc[1] < d->ny - 1 ? bc[1] : d->ny - 1) : 0;
    } else {
        p.rlo = 1; p.rhi = d->nx - 1; p.clo = 1; p.chi = d->ny - 1;
    }
    p.nsrc = d->nsrc;
    for (int k = 0; k < d->nsrc; ++k) { p.src_i[k] = d->src_i[k]; p.src_j[k] = d->src_j[k]; }
    p.src_wave = d->src_wave; p.wave_len = d->wave_len;
    p.nprobe = d->nprobe;
    for (int k = 0; k < d->nprobe; ++k) { p.probe_i[k] = d->probe_i[k]; p.probe_j[k] = d->probe_j[k]; }
    p.probe_out = d->probe_out; p.probe_len = d->probe_len;
    float *A[6] = {d->dz, d->iz, d->hx, d->hy, d->ihx, d->ihy};
    float *B[6] = {d->dz2, d->iz2, d->hx2, d->hy2, d->ihx2, d->ihy2};
    const int L = pl.launches;
    // free-space tiles: the first launch (the deepest: its tiles read every
    // row a later launch's tiles read) records them, the others use them; a
    // fresh tag per recording keeps flags of other calls from matching,
    // unless the caller promises an unchanged medium (tile_keep) and the
    // flags were recorded with this workspace, arrays and tile geometry at
    // a depth at least this call's
    const bool vac = d->tile_work && !pl.mw && L >= 2 && d->tile_elems >= tm_tiles(d, pl);
    bool reuse = false;
    if (vac) {
        const int kmax = nfused / L + (nfused % L ? 1 : 0);
        TileRecord now{d->tile_work, d->md.naz, d->md.nbz, d->iz, d->iz2, d->nx, d->ny, pl.rows,
                       pl.cols, pl.halo, kmax, 0u};
        std::lock_guard<std::mutex> g(tile_records_mutex());
        auto &recs = tile_records();
        auto it = std::find_if(recs.begin(), recs.end(),
                               [&](const TileRecord &r) { return r.work == d->tile_work; });
        if (d->tile_keep && it != recs.end() && it->same_tiles(now) && it->depth >= kmax) {
            reuse = true;
            p.vac_tag = it->tag;
        } else {
            static std::atomic<unsigned> calls{0};
            now.tag = (calls.fetch_add(1) + 1u) << 1;
            p.vac_tag = now.tag;
            if (it != recs.end()) {
                *it = now;
            } else {
                // a bounded list: a workspace whose record was dropped only
                // records again on its next call
                if (recs.size() >= 256) recs.erase(recs.begin());
                recs.push_back(now);
            }
        }
        p.tile_vac = d->tile_work;
    }
    int done = 0;
    for (int l = 0; l < L; ++l) {
        const int k = nfused / L + (l < nfused % L ? 1 : 0);
        p.vac_mode = vac ? ((l == 0 && !reuse) ? 1 : 2) : 0;
        float **in = (l % 2 == 0) ? A : B;
        float **out = (l % 2 == 0) ? B : A;
        p.dz_i = in[0]; p.iz_i = in[1]; p.hx_i = in[2]; p.hy_i = in[3]; p.ihx_i = in[4]; p.ihy_i = in[5];
        p.dz_o = out[0]; p.iz_o = out[1]; p.hx_o = out[2]; p.hy_o = out[3]; p.ihx_o = out[4]; p.ihy_o = out[5];
        p.ez_o = (l == L - 1 && !pl.tail_unfused) ? d->ez : nullptr;
        p.t_first = t0 + done + 1;
        if (pl.mw) rc = launch_fd2d_mw(p, k, lossy, pml, pl.rows, pl.frows, S(stream));
        else rc = launch_fd2d_tb(p, k, pl.cols, lossy, pml, TB_WHOLE, S(stream));
        if (rc) return rc;
        done += k;
    }
    if (pl.tail_unfused) {
        const int t = t0 + done + 1;
        SrcList src{};
        src.n = d->nsrc;
        for (int k = 0; k < d->nsrc; ++k) { src.i[k] = d->src_i[k]; src.j[k] = d->src_j[k]; }
        src.wave = d->src_wave; src.wave_len = d->wave_len; src.t = t;
        PmlPtrs pp = to_ptrs(d->pml);
        rc = launch_fd2d_dfield(d->nx, d->ny, pml ? &pp : nullptr, d->dz, d->hx, d->hy, src, S(stream));
        if (rc) return rc;
        rc = launch_fd2d_efield(d->nx, d->ny, d->md.naz, d->md.nbz, d->dz, d->iz, d->ez, S(stream));
        if (rc) return rc;
        rc = launch_fd2d_hfield(d->nx, d->ny, pml ? &pp : nullptr, d->ez, d->ihx, d->ihy, d->hx,
                                d->hy, S(stream));
        if (rc) return rc;
        if (d->nprobe) {
            rc = launch_fd2d_probe(d->ny, d->ez, d->nprobe, d->probe_i, d->probe_j, d->probe_out,
                                   d->probe_len, t - 1, S(stream));
            if (rc) return rc;
        }
    }
    return 0;
}

// ------------------------------------------------ 1-D D/E form (lossy, Debye)

static int medium1d_check(const char *fn, const fdtd_medium1d *md)
{
    if (!md || !md->nax || !md->nbx) return fail("%s: md->nax and md->nbx are required", fn);
    if ((md->ncx != nullptr) != (md->ndx != nullptr)) return fail("%s: ncx and ndx go together", fn);
    return 0;
}

int fd1d_dxfield(int t, int nx, float *dx, const float *hy, int src_pos, float src_val,
                 void *stream)
{
    (void)t;
    clear_error();
    if (nx < 2) return fail("fd1d_dxfield: nx=%d", nx);
    if (!dx || !hy) return fail("fd1d_dxfield: dx and hy are required");
    if (src_pos >= nx) return fail("fd1d_dxfield: src_pos=%d outside the grid", src_pos);
    return launch_dxfield(nx, dx, hy, src_pos, src_val, src_pos >= 1, S(stream));
}

int fd1d_exfield_dform(int nx, const fdtd_medium1d *md, const float *dx, float *ix, float *sx,
                       float *ex, void *stream)
{
    clear_error();
    if (nx < 2) return fail("fd1d_exfield_dform: nx=%d", nx);
    int rc = medium1d_check("fd1d_exfield_dform", md);
    if (rc) return rc;
    if (!dx || !ix || !ex || (md->ncx && !sx)) return fail("fd1d_exfield_dform: missing field");
    return launch_exfield_dform(nx, *md, dx, ix, sx, ex, S(stream));
}

int fd1d_fourier(int t, int nf, int nx, float dt, const float *freq, const float *ex,
                 const fdtd_ftrans *ft, void *stream)
{
    clear_error();
    if (nx < 11 || !ex) return fail("fd1d_fourier: nx=%d (probe ex[10]), ex required", nx);
    if (nf < 1 || nf > FDTD_MAX_FREQ || !freq) return fail("fd1d_fourier: nf=%d, freq (host)", nf);
    if (!ft || !ft->r_pt || !ft->i_pt) return fail("fd1d_fourier: ft->r_pt and ft->i_pt are required");
    return launch_fourier1d(t, nf, nx, dt, freq, ex, 10, *ft, S(stream));
}

int fd1d_dielectric_dform(int nx, float dt, float chi, float tau, float epsr, float sigma,
                          const fdtd_medium1d *md, void *stream)
{
    clear_error();
    if (nx < 2) return fail("fd1d_dielectric_dform: nx=%d", nx);
    int rc = medium1d_check("fd1d_dielectric_dform", md);
    if (rc) return rc;
    if (md->ncx && !(tau > 0.0f)) return fail("fd1d_dielectric_dform: tau=%g", tau);
    return launch_debye(nx, dt, chi, tau, epsr, sigma, *md, S(stream));
}

size_t fd1d_dform_work_elems(int nx)
{
    return nx < 1 ? 0 : 5 * (size_t)nx + 4;
}

// The fused D/E-form run: an even number L of launches of up to DF_MAXK
// steps, ping-ponging between the caller's state set and the partner set in
// work (so the result ends in the caller's arrays).
static int dform_fused(const fd1d_dform_desc *d, int t0, int nsteps, hipStream_t s)
{
    const int nx = d->nx;
    int L = (nsteps + DF_MAXK - 1) / DF_MAXK;
    L += L % 2;
    float *w = d->work;
    float *A[6] = {d->dx, d->ix, d->sx, d->ex, d->hy, d->bc};
    float *B[6] = {w, w + nx, w + 2 * (size_t)nx, w + 3 * (size_t)nx, w + 4 * (size_t)nx,
                   w + 5 * (size_t)nx};
    int done = 0;
    for (int l = 0; l < L; ++l) {
        const int k = nsteps / L + (l < nsteps % L ? 1 : 0);
        float **in = (l % 2 == 0) ? A : B;
        float **out = (l % 2 == 0) ? B : A;
        DformTbParams p{};
        p.nx = nx; p.t_first = t0 + done + 1; p.K = k;
        p.dx_i = in[0]; p.ix_i = in[1]; p.sx_i = in[2]; p.ex_i = in[3]; p.hy_i = in[4]; p.bc_i = in[5];
        p.dx_o = out[0]; p.ix_o = out[1]; p.sx_o = out[2]; p.ex_o = out[3]; p.hy_o = out[4]; p.bc_o = out[5];
        p.nax = d->md.nax; p.nbx = d->md.nbx; p.ncx = d->md.ncx; p.ndx = d->md.ndx;
        p.src_pos = d->src_pos >= 1 ? d->src_pos : -1;
        p.nf = d->nf;
        p.r_pt = d->ft.r_pt; p.i_pt = d->ft.i_pt; p.r_in = d->ft.r_in; p.i_in = d->ft.i_in;
        for (int q = 0; q < k; ++q) {
            const int t = p.t_first + q;
            p.src_val[q] = p.src_pos >= 1 ? d->src_wave[t - 1] : 0.0f;
            for (int f = 0; f < d->nf; ++f) dft_twiddle(t, d->freq[f], d->dt, &p.tw_c[q][f], &p.tw_s[q][f]);
        }
        p.nprobe = d->nprobe;
        for (int q = 0; q < d->nprobe; ++q) p.probe_k[q] = d->probe_k[q];
        p.probe_out = d->probe_out; p.probe_len = d->probe_len;
        int rc = launch_dform_tb(p, s);
        if (rc) return rc;
        done += k;
    }
    return 0;
}

int fd1d_dform_run(const fd1d_dform_desc *d, int t0, int nsteps, void *stream)
{
    clear_error();
    if (!d) return fail("fd1d_dform_run: NULL descriptor");
    const int nx = d->nx;
    if (nx < 11) return fail("fd1d_dform_run: nx=%d", nx);
    int rc = medium1d_check("fd1d_dform_run", &d->md);
    if (rc) return rc;
    if (!d->dx || !d->ix || !d->ex || !d->hy || !d->bc || (d->md.ncx && !d->sx))
        return fail("fd1d_dform_run: dx, ix, ex, hy, bc (and sx with a Debye medium) are required");
    if (t0 < 0 || nsteps < 0) return fail("fd1d_dform_run: t0=%d nsteps=%d", t0, nsteps);
    if (d->src_pos >= nx) return fail("fd1d_dform_run: src_pos=%d", d->src_pos);
    if (d->src_pos >= 1 && (!d->src_wave || d->wave_len < t0 + nsteps))
        return fail("fd1d_dform_run: src_wave (host) must cover step %d", t0 + nsteps);
    if (d->nf) {
        if (d->nf < 0 || d->nf > FDTD_MAX_FREQ || !d->freq) return fail("fd1d_dform_run: nf=%d", d->nf);
        if (!d->ft.r_pt || !d->ft.i_pt) return fail("fd1d_dform_run: DFT planes required");
    }
    if (d->nprobe < 0 || d->nprobe > FDTD_MAX_PROBES) return fail("fd1d_dform_run: nprobe=%d", d->nprobe);
    if (d->nprobe) {
        if (!d->probe_out || d->probe_len < t0 + nsteps) return fail("fd1d_dform_run: probe_out");
        for (int p = 0; p < d->nprobe; ++p)
            if (d->probe_k[p] < 0 || d->probe_k[p] >= nx) return fail("fd1d_dform_run: probe %d", p);
    }
    hipStream_t s = S(stream);
    if (d->work && nsteps >= 2) {
        if (d->work_elems < fd1d_dform_work_elems(nx))
            return fail("fd1d_dform_run: work has %zu floats, %zu needed", d->work_elems,
                        fd1d_dform_work_elems(nx));
        if (d->nf > DF_NF)
            return fail("fd1d_dform_run: the fused path takes at most %d frequencies (pass work = "
                        "NULL for the per-step path)", DF_NF);
        return dform_fused(d, t0, nsteps, s);
    }
    int zeros[FDTD_MAX_PROBES] = {0};
    for (int k = 0; k < nsteps; ++k) {
        const int t = t0 + k + 1;
        const bool src = d->src_pos >= 1;
        rc = launch_dxfield(nx, d->dx, d->hy, d->src_pos, src ? d->src_wave[t - 1] : 0.0f, src, s);
        if (!rc) rc = launch_exfield_dform(nx, d->md, d->dx, d->ix, d->sx, d->ex, s);
        if (!rc && d->nf) rc = launch_fourier1d(t, d->nf, nx, d->dt, d->freq, d->ex, 10, d->ft, s);
        if (!rc) rc = launch_fd1d_hyfield<float>(nx, d->ex, d->hy, d->bc, s);
        if (!rc && d->nprobe)
            rc = launch_fd2d_probe(nx, d->ex, d->nprobe, zeros, d->probe_k, d->probe_out,
                                   d->probe_len, t - 1, s);
        if (rc) return rc;
    }
    return 0;
}

// ------------------------------------------------ plane wave (TF/SF) and DFT

static int tfsf_box_check(const char *fn, int nx, int ny, int npml)
{
    if (npml < 2 || npml - 1 > nx - npml || npml - 1 > ny - npml)
        return fail("%s: npml=%d gives no TF/SF box in %dx%d", fn, npml, nx, ny);
    return 0;
}

int fd2d_ezinct(int ny, float *ezi, const float *hxi, float *bc, void *stream)
{
    clear_error();
    if (ny < 3) return fail("fd2d_ezinct: ny=%d", ny);
    if (!ezi || !hxi || !bc) return fail("fd2d_ezinct: ezi, hxi, bc are required");
    return launch_ezinct(ny, ezi, hxi, bc, 0, 0, 0.0f, S(stream));
}

int fd2d_dfield_inc(int t, int nx, int ny, const fdtd_pml *pml, float *ezi, float *dz,
                    const float *hx, const float *hy, int inc_pos, float inc_val, void *stream)
{
    clear_error();
    if (!ezi) return fail("fd2d_dfield_inc: ezi is required");
    if (inc_pos < 0 || inc_pos >= ny) return fail("fd2d_dfield_inc: inc_pos=%d outside [0,%d)", inc_pos, ny);
    int rc = fd2d_dfield(t, nx, ny, pml, dz, hx, hy, 0, nullptr, nullptr, nullptr, stream);
    if (rc) return rc;
    return launch_set(ezi, inc_pos, inc_val, S(stream));
}

int fd2d_inctdz(int nx, int ny, int npml, const float *hxi, float *dz, void *stream)
{
    clear_error();
    int rc = tfsf_box_check("fd2d_inctdz", nx, ny, npml);
    if (rc) return rc;
    if (!hxi || !dz) return fail("fd2d_inctdz: hxi and dz are required");
    return launch_inctdz(nx, ny, npml, hxi, dz, S(stream));
}

int fd2d_hxinct(int ny, const float *ezi, float *hxi, void *stream)
{
    clear_error();
    if (ny < 2) return fail("fd2d_hxinct: ny=%d", ny);
    if (!ezi || !hxi) return fail("fd2d_hxinct: ezi and hxi are required");
    return launch_hxinct(ny, ezi, hxi, S(stream));
}

int fd2d_incthx(int nx, int ny, int npml, const float *ezi, float *hx, void *stream)
{
    clear_error();
    int rc = tfsf_box_check("fd2d_incthx", nx, ny, npml);
    if (rc) return rc;
    if (!ezi || !hx) return fail("fd2d_incthx: ezi and hx are required");
    return launch_incthx(nx, ny, npml, ezi, hx, S(stream));
}

int fd2d_incthy(int nx, int ny, int npml, const float *ezi, float *hy, void *stream)
{
    clear_error();
    int rc = tfsf_box_check("fd2d_incthy", nx, ny, npml);
    if (rc) return rc;
    if (!ezi || !hy) return fail("fd2d_incthy: ezi and hy are required");
    return launch_incthy(nx, ny, npml, ezi, hy, S(stream));
}

static int fourier_check(const char *fn, int nf, int ny, const float *freq, const float *ezi,
                         int probe, const fdtd_ftrans *ft)
{
    if (nf < 1 || nf > FDTD_MAX_FREQ) return fail("%s: nf=%d outside [1,%d]", fn, nf, FDTD_MAX_FREQ);
    if (!freq) return fail("%s: freq (host array) is required", fn);
    if (!ft || !ft->r_pt || !ft->i_pt) return fail("%s: ft->r_pt and ft->i_pt are required", fn);
    if ((ft->r_in != nullptr) != (ft->i_in != nullptr)) return fail("%s: r_in and i_in go together", fn);
    if (ft->r_in && (!ezi || probe < 0 || probe >= ny))
        return fail("%s: the incident transform needs ezi with %d cells", fn, probe + 1);
    return 0;
}

int fd2d_fourier(int t, int nf, int nx, int ny, float dt, const float *freq, const float *ezi,
                 const float *ez, const fdtd_ftrans *ft, void *stream)
{
    clear_error();
    if (nx < 1 || ny < 1 || !ez) return fail("fd2d_fourier: grid %dx%d, ez required", nx, ny);
    int rc = fourier_check("fd2d_fourier", nf, ny, freq, ezi, 6, ft);
    if (rc) return rc;
    return launch_fourier(t, nf, nx, ny, dt, freq, ezi, 6, ez, *ft, S(stream));
}

size_t fd2d_tfsf_work_elems(int ny, int npml)
{
    if (npml < 2 || ny - 2 * npml + 2 < 1) return 0;
    return (size_t)INC_CHUNK * (size_t)(ny - 2 * npml + 4);
}

// The fused plane-wave run: per chunk of up to INC_CHUNK steps, the incident
// grid advances in one workgroup (inc_chunk_kernel, writing the table the
// 2-D launches read), then the temporally blocked launches apply the TF/SF
// corrections and accumulate the DFT (as fd2d_tm_run otherwise).
static int tfsf_fused(const fd2d_tm_desc *d, const fd2d_tfsf_desc *b, int t0, int nsteps,
                      hipStream_t s0)
{
    const bool lossy = d->md.nbz != nullptr;
    const bool pml = pml_all(&d->pml);
    const int stride = d->ny - 2 * b->npml + 4;
    int rc = 0;
    int done = 0;
    // with the DFT a level carries 6 more registers per column: default depth
    // 4 and short row blocks (more waves; MI355X sweeps at 1024^2, npml 80, 3
    // frequencies: 38k Mcell-updates/s at depth 8 rows 16, 60k at depth 4
    // rows 4)
    fd2d_tm_desc dd = *d;
    dd.form = FDTD_TM_FORM_STRIP;          // TF/SF corrections and DFT: strip form
    if (b->nf && !dd.depth) dd.depth = 4;
    // the plane-wave launches run the instantiation that carries the DFT
    // planes, whose level loop stays unrolled up to depth 8 only (past it the
    // twiddle table is indexed at run time and the launch parameters go to
    // scratch memory); the depth changes no result
    if (!dd.depth) {
        int h;
        tm_defaults(d->nx, &dd.depth, &h);
    }
    if (dd.depth > 8) dd.depth = 8;
    if (b->nf && !dd.rows) {
        const int h = d->nx / 220;
        dd.rows = h < 4 ? 4 : (h > 512 ? 512 : h);
    }
    d = &dd;
    while (done < nsteps) {
        const int n = nsteps - done < INC_CHUNK ? nsteps - done : INC_CHUNK;
        const int tc = t0 + done;                  // steps tc+1 .. tc+n
        Plan pl;
        rc = plan_tm(d, n, &pl);
        if (rc) return rc;
        if (pl.tail_unfused || pl.launches == 0) {
            // too short for an even number of launches: the per-step kernels
            fd2d_tfsf_desc one = *b;
            one.work = nullptr;
            rc = fd2d_tfsf_run(d, &one, tc, n, s0);
            if (rc) return rc;
            done += n;
            continue;
        }
        IncChunk ic{};
        ic.n = n; ic.npml = b->npml; ic.stride = stride; ic.inc_pos = b->inc_pos;
        ic.probe = b->dft_probe; ic.nf = b->nf;
        for (int k = 0; k < n; ++k) {
            ic.wave[k] = b->inc_wave[tc + k];
            for (int f = 0; f < b->nf; ++f) dft_twiddle(tc + k + 1, b->freq[f], b->dt, &ic.c[k][f], &ic.s[k][f]);
        }
        rc = launch_inc_chunk(d->ny, b->ezi, b->hxi, b->bc, b->nf ? b->ft.r_in : nullptr,
                              b->nf ? b->ft.i_in : nullptr, b->work, ic, s0);
        if (rc) return rc;
        TbParams p{};
        p.nx = d->nx; p.ny = d->ny;
        p.naz = d->md.naz; p.nbz = d->md.nbz;
        if (pml) p.pml = to_ptrs(d->pml);
        p.halo = pl.halo; p.rows = pl.rows;
        if (pml) {
            const int *br = d->band_rows, *bc = d->band_cols;
            p.rlo = br[1] > br[0] ? (br[0] > 1 ? br[0] : 1) : 0;
            p.rhi = br[1] > br[0] ? (br[1] < d->nx - 1 ? br[1] : d->nx - 1) : 0;
            p.clo = bc[1] > bc[0] ? (bc[0] > 1 ? bc[0] : 1) : 0;
            p.chi = bc[1] > bc[0] ? (bc[1] < d->ny - 1 ? bc[1] : d->ny - 1) : 0;
        } else {
            p.rlo = 1; p.rhi = d->nx - 1; p.clo = 1; p.chi = d->ny - 1;
        }
        p.nsrc = d->nsrc;
        for (int k = 0; k < d->nsrc; ++k) { p.src_i[k] = d->src_i[k]; p.src_j[k] = d->src_j[k]; }
        p.src_wave = d->src_wave; p.wave_len = d->wave_len;
        p.nprobe = d->nprobe;
        for (int k = 0; k < d->nprobe; ++k) { p.probe_i[k] = d->probe_i[k]; p.probe_j[k] = d->probe_j[k]; }
        p.probe_out = d->probe_out; p.probe_len = d->probe_len;
        p.tf_npml = b->npml; p.inc = b->work; p.inc_t0 = tc; p.inc_stride = stride;
        p.nf = b->nf; p.ft_r = b->ft.r_pt; p.ft_i = b->ft.i_pt;
        float *A[6] = {d->dz, d->iz, d->hx, d->hy, d->ihx, d->ihy};
        float *B[6] = {d->dz2, d->iz2, d->hx2, d->hy2, d->ihx2, d->ihy2};
        const int L = pl.launches;
        int dl = 0;
        for (int l = 0; l < L; ++l) {
            const int k = n / L + (l < n % L ? 1 : 0);
            float **in = (l % 2 == 0) ? A : B;
            float **out = (l % 2 == 0) ? B : A;
            p.dz_i = in[0]; p.iz_i = in[1]; p.hx_i = in[2]; p.hy_i = in[3]; p.ihx_i = in[4]; p.ihy_i = in[5];
            p.dz_o = out[0]; p.iz_o = out[1]; p.hx_o = out[2]; p.hy_o = out[3]; p.ihx_o = out[4]; p.ihy_o = out[5];
            p.ez_o = (l == L - 1) ? d->ez : nullptr;
            p.t_first = tc + dl + 1;
            for (int q = 0; q < k; ++q)
                for (int f = 0; f < b->nf; ++f) dft_twiddle(p.t_first + q, b->freq[f], b->dt, &p.tw_c[q][f], &p.tw_s[q][f]);
            rc = launch_fd2d_tb(p, k, pl.cols, lossy, pml, TB_WHOLE, s0);
            if (rc) return rc;
            dl += k;
        }
        done += n;
    }
    return 0;
}

int fd2d_tfsf_run(const fd2d_tm_desc *d, const fd2d_tfsf_desc *b, int t0, int nsteps, void *stream)
{
    clear_error();
    int rc = tm_validate(d, t0, nsteps);
    if (rc) return rc;
    if (!b) return fail("fd2d_tfsf_run: NULL plane-wave descriptor");
    rc = tfsf_box_check("fd2d_tfsf_run", d->nx, d->ny, b->npml);
    if (rc) return rc;
    if (!b->ezi || !b->hxi || !b->bc) return fail("fd2d_tfsf_run: ezi, hxi, bc are required");
    if (b->inc_pos < 0 || b->inc_pos >= d->ny) return fail("fd2d_tfsf_run: inc_pos=%d", b->inc_pos);
    if (!b->inc_wave || b->inc_len < t0 + nsteps)
        return fail("fd2d_tfsf_run: inc_wave (host) must cover step %d", t0 + nsteps);
    if (b->nf) {
        rc = fourier_check("fd2d_tfsf_run", b->nf, d->ny, b->freq, b->ezi, b->dft_probe, &b->ft);
        if (rc) return rc;
    }
    const bool pml = pml_all(&d->pml);
    if (b->work) {
        if (b->work_elems < fd2d_tfsf_work_elems(d->ny, b->npml))
            return fail("fd2d_tfsf_run: work has %zu floats, %zu needed", b->work_elems,
                        fd2d_tfsf_work_elems(d->ny, b->npml));
        if (b->nf > TB_MAX_FREQ || !pml)
            return fail("fd2d_tfsf_run: the fused path takes the PML form and at most %d "
                        "frequencies (pass work = NULL for the per-step path)", TB_MAX_FREQ);
        if (!d->dz2 || !d->hx2 || !d->hy2 || (d->md.nbz && !d->iz2) || (pml && (!d->ihx2 || !d->ihy2)))
            return fail("fd2d_tfsf_run: the fused path needs the partner state set");
        return tfsf_fused(d, b, t0, nsteps, S(stream));
    }
    PmlPtrs pp = to_ptrs(d->pml);
    hipStream_t s = S(stream);
    for (int k = 0; k < nsteps; ++k) {
        const int t = t0 + k + 1;
        SrcList src{};
        src.n = d->nsrc;
        for (int q = 0; q < d->nsrc; ++q) { src.i[q] = d->src_i[q]; src.j[q] = d->src_j[q]; }
        src.wave = d->src_wave; src.wave_len = d->wave_len; src.t = t;
        rc = launch_ezinct(d->ny, b->ezi, b->hxi, b->bc, 1, b->inc_pos, b->inc_wave[t - 1], s);
        if (!rc) rc = launch_fd2d_dfield(d->nx, d->ny, pml ? &pp : nullptr, d->dz, d->hx, d->hy, src, s);
        if (!rc) rc = launch_inctdz(d->nx, d->ny, b->npml, b->hxi, d->dz, s);
        if (!rc) rc = launch_fd2d_efield(d->nx, d->ny, d->md.naz, d->md.nbz, d->dz, d->iz, d->ez, s);
        if (!rc && b->nf)
            rc = launch_fourier(t, b->nf, d->nx, d->ny, b->dt, b->freq, b->ezi, b->dft_probe, d->ez,
                                b->ft, s);
        if (!rc) rc = launch_hxinct(d->ny, b->ezi, b->hxi, s);
        if (!rc) rc = launch_fd2d_hfield(d->nx, d->ny, pml ? &pp : nullptr, d->ez, d->ihx, d->ihy,
                                         d->hx, d->hy, s);
        if (!rc) rc = launch_incthx(d->nx, d->ny, b->npml, b->ezi, d->hx, s);
        if (!rc) rc = launch_incthy(d->nx, d->ny, b->npml, b->ezi, d->hy, s);
        if (!rc && d->nprobe)
            rc = launch_fd2d_probe(d->ny, d->ez, d->nprobe, d->probe_i, d->probe_j, d->probe_out,
                                   d->probe_len, t - 1, s);
        if (rc) return rc;
    }
    return 0;
}

// ------------------------------------------------ learned boundary

size_t lbc_hist_elems(int nx, int ny)
{
    if (nx < 8 || ny < 8) return 0;
    return (size_t)lbc::npos(nx, ny) * lbc::HIST * lbc::NS;
}

// The learned run's work buffer: the frame kernel's geometry tables, then
// the free-space flags of the interior part's tiles (one unsigned per tile;
// room for the smallest tiles a plan allows: strips of 8 columns, blocks of
// one row).
static size_t lbc_tile_offset(int nx, int ny)
{
    return (lbc_frame_table_offset(nx, ny, LBC_MAX_DEPTH + 1) + 255) / 256 * 256;
}

static size_t lbc_tile_capacity(int nx, int ny) { return (size_t)nx * (size_t)((ny + 7) / 8); }

size_t lbc_work_bytes(int nx, int ny)
{
    if (nx < 64 || ny < 64) return 0;
    return lbc_tile_offset(nx, ny) + lbc_tile_capacity(nx, ny) * sizeof(unsigned);
}

int fd2d_lbc_margin(void) { return FDTD_LBC_MARGIN; }

int fd2d_lbc_run(const fd2d_tm_desc *d, const fd2d_lbc_desc *b, int t0, int nsteps, void *stream)
{
    clear_error();
    int rc = tm_validate(d, t0, nsteps);
    if (rc) return rc;
    if (!pml_none(&d->pml)) return fail("fd2d_lbc_run: the learned boundary replaces the pml (pass none)");
    if (!b || !b->params || !b->wbuf || !b->hist || !b->hist2 || !b->work)
        return fail("fd2d_lbc_run: params, wbuf, hist, hist2 and work are required");
    if (d->nx < 64 || d->ny < 64) return fail("fd2d_lbc_run: grid %dx%d below 64x64", d->nx, d->ny);
    Plan pl;
    fd2d_tm_desc ds = *d;
    ds.form = FDTD_TM_FORM_STRIP;         // the split launches are single-wave strips
    if (!ds.depth) ds.depth = 8;          // and fuse at most LBC_MAX_DEPTH steps (the frame's tables)
    rc = plan_tm(&ds, 2 * LBC_MAX_DEPTH, &pl);
    if (rc) return rc;
    if (pl.depth > LBC_MAX_DEPTH)
        return fail("fd2d_lbc_run: depth %d above %d", pl.depth, LBC_MAX_DEPTH);
    // sources and probes outside the frame kernel's regions (which apply no
    // source and record no probe): 3K+6 deep at the corner chunks
    const int M = std::max(FDTD_LBC_MARGIN, 3 * pl.depth + 6);
    for (int k = 0; k < d->nsrc; ++k)
        if (d->src_i[k] < M || d->src_i[k] >= d->nx - M || d->src_j[k] < M || d->src_j[k] >= d->ny - M)
            return fail("fd2d_lbc_run: source %d within %d cells of an edge", k, M);
    for (int k = 0; k < d->nprobe; ++k)
        if (d->probe_i[k] < M || d->probe_i[k] >= d->nx - M || d->probe_j[k] < M ||
            d->probe_j[k] >= d->ny - M)
            return fail("fd2d_lbc_run: probe %d within %d cells of an edge", k, M);
    if (nsteps == 0) return 0;
    if (nsteps < 2) return fail("fd2d_lbc_run: nsteps=%d (at least 2: launches come in pairs)", nsteps);
    const bool lossy = d->md.nbz != nullptr;
    if (!d->dz2 || !d->hx2 || !d->hy2 || (lossy && !d->iz2))
        return fail("fd2d_lbc_run: the partner state set (dz2, iz2, hx2, hy2) is required");
    int L = (nsteps + pl.depth - 1) / pl.depth;
    L += L % 2;
    TbParams p{};
    p.nx = d->nx; p.ny = d->ny;
    p.naz = d->md.naz; p.nbz = d->md.nbz;
    // this run's launches fuse up to pl.depth steps (the plan above was made
    // only to fix depth, rows and columns)
    p.halo = ((pl.depth + pl.cols - 1) / pl.cols) * pl.cols;
    p.rows = pl.rows;
    // interior part: waves whose reads and writes stay DZ or more cells from
    // every edge.  The frame kernel reads less than DZ = 3K+6 cells deep and
    // writes less than K+3 deep, so the interior part of launch l+1 needs
    // neither the frame of launch l nor conflicts with it.
    const int DZ = 3 * pl.depth + 6;
    p.rlo = DZ; p.rhi = d->nx - DZ; p.clo = DZ; p.chi = d->ny - DZ;
    // no PML: boundary-part tiles clear of the edge rows and columns take
    // the interior body
    p.erlo = 1; p.erhi = d->nx - 1; p.eclo = 1; p.echi = d->ny - 1;
    p.nsrc = d->nsrc;
    for (int k = 0; k < d->nsrc; ++k) { p.src_i[k] = d->src_i[k]; p.src_j[k] = d->src_j[k]; }
    p.src_wave = d->src_wave; p.wave_len = d->wave_len;
    p.nprobe = d->nprobe;
    for (int k = 0; k < d->nprobe; ++k) { p.probe_i[k] = d->probe_i[k]; p.probe_j[k] = d->probe_j[k]; }
    p.probe_out = d->probe_out; p.probe_len = d->probe_len;
    float *A[4] = {d->dz, d->iz, d->hx, d->hy};
    float *B[4] = {d->dz2, d->iz2, d->hx2, d->hy2};
    float *H[2] = {b->hist, b->hist2};
    Aux *ax = nullptr;
    rc = aux_for_current_device(&ax);
    if (rc) return rc;
    // geometry tables of the (at most two) launch depths, built into the
    // caller's work buffer on the frame stream, ahead of the frame kernels
    const int kq = nsteps / L, kr = nsteps % L;
    char *wk = static_cast<char *>(b->work);
    void *tabq = wk + lbc_frame_table_offset(d->nx, d->ny, kq);
    void *tabr = kr ? wk + lbc_frame_table_offset(d->nx, d->ny, kq + 1) : nullptr;
    // three streams: interior part P0 on the caller's, boundary part P1 on
    // ax->s, frame F on ax->s2 (one merged launch would start ahead of the
    // frame and keep it from the CUs).  Launch l:
    //   P0(l) after P0(l-1), P1(l-1)   (reads and overwrites what they used)
    //   P1(l) after P0(l-1), P1(l-1), F(l-1)
    //   F(l)  after P1(l-1), F(l-1)    (P0 stays DZ cells clear of the frame)
    // events: join0 = last P0, join = last P1, join2 = last F.
    hipStream_t s0 = S(stream);
    rc = hip_check(hipEventRecord(ax->fork, s0), "fork event");
    if (!rc) rc = hip_check(hipStreamWaitEvent(ax->s, ax->fork, 0), "fork wait");
    if (!rc) rc = hip_check(hipStreamWaitEvent(ax->s2, ax->fork, 0), "fork wait");
    if (!rc) rc = launch_lbc_frame_setup(d->nx, d->ny, kq, tabq, ax->s2);
    if (!rc && kr) rc = launch_lbc_frame_setup(d->nx, d->ny, kq + 1, tabr, ax->s2);
    if (rc) return rc;
    // free-space tiles of the interior part, as in fd2d_tm_run: the first
    // launch (the deepest) records them with a tag fresh to the call, the
    // later ones run them without naz, nbz, iz (iz is +0 in both sets there)
    {
        const int w = 64 * pl.cols - 2 * p.halo;
        const size_t tiles = (size_t)((d->ny + w - 1) / w) * (size_t)((d->nx + pl.rows - 1) / pl.rows);
        if (tiles <= lbc_tile_capacity(d->nx, d->ny)) {
            static std::atomic<unsigned> calls{0};
            p.tile_vac = reinterpret_cast<unsigned *>(wk + lbc_tile_offset(d->nx, d->ny));
            p.vac_tag = (calls.fetch_add(1) + 1u) << 1;
        }
    }
    int done = 0;
    for (int l = 0; l < L; ++l) {
        const int k = kq + (l < kr ? 1 : 0);
        float **in = (l % 2 == 0) ? A : B;
        float **out = (l % 2 == 0) ? B : A;
        p.vac_mode = p.tile_vac ? (l == 0 ? 1 : 2) : 0;
        p.dz_i = in[0]; p.iz_i = in[1]; p.hx_i = in[2]; p.hy_i = in[3];
        p.ihx_i = p.ihy_i = nullptr;
        p.dz_o = out[0]; p.iz_o = out[1]; p.hx_o = out[2]; p.hy_o = out[3];
        p.ihx_o = p.ihy_o = nullptr;
        p.ez_o = (l == L - 1) ? d->ez : nullptr;
        p.t_first = t0 + done + 1;
        p.skip_w = k + 3;          // the frame kernel writes the outer k+3 cells
        LbcFrameParams f{};
        f.nx = d->nx; f.ny = d->ny; f.K = k;
        f.dz_i = in[0]; f.iz_i = lossy ? in[1] : nullptr; f.hx_i = in[2]; f.hy_i = in[3];
        f.dz_o = out[0]; f.iz_o = lossy ? out[1] : nullptr; f.hx_o = out[2]; f.hy_o = out[3];
        f.ez_o = p.ez_o;
        f.naz = d->md.naz; f.nbz = d->md.nbz;
        f.hist_in = H[l % 2]; f.hist_out = H[(l + 1) % 2];
        f.w = b->wbuf; f.params = b->params;
        f.tab = (l < kr) ? tabr : tabq;
        if (l > 0) {
            // waits on the events of launch l-1 (recorded below), issued
            // before this launch records them again
            rc = hip_check(hipStreamWaitEvent(ax->s, ax->join0, 0), "P1 wait P0");
            if (!rc) rc = hip_check(hipStreamWaitEvent(ax->s, ax->join2, 0), "P1 wait F");
            if (!rc) rc = hip_check(hipStreamWaitEvent(ax->s2, ax->join, 0), "F wait P1");
            if (!rc) rc = hip_check(hipStreamWaitEvent(s0, ax->join, 0), "P0 wait P1");
            if (rc) return rc;
        }
        rc = launch_lbc_frame(f, ax->s2);
        if (!rc) rc = hip_check(hipEventRecord(ax->join2, ax->s2), "F event");
        if (!rc) rc = launch_fd2d_tb(p, k, pl.cols, lossy, false, TB_BOUNDARY, ax->s);
        if (!rc) rc = hip_check(hipEventRecord(ax->join, ax->s), "P1 event");
        if (!rc) rc = launch_fd2d_tb(p, k, pl.cols, lossy, false, TB_INTERIOR, s0);
        if (!rc) rc = hip_check(hipEventRecord(ax->join0, s0), "P0 event");
        if (rc) return rc;
        done += k;
    }
    rc = hip_check(hipStreamWaitEvent(s0, ax->join, 0), "join wait");
    if (!rc) rc = hip_check(hipStreamWaitEvent(s0, ax->join2, 0), "join wait");
    if (rc) return rc;
    return 0;
}

// ------------------------------------------------------------ batches

int fd2d_batch_units(const fd2d_grid *grids, int ngrids, int depth, int rows, int cols,
                     int *nunits)
{
    clear_error();
    if (!grids || ngrids <= 0 || !nunits) return fail("fd2d_batch_units: grids/ngrids/nunits");
    BatchGeom g;
    int rc = batch_geom(depth, rows, cols, &g);
    if (rc) return rc;
    const int w = 64 * g.cols - 2 * g.halo;
    long long n = 0;
    for (int k = 0; k < ngrids; ++k) {
        if (grids[k].nx < 2 || grids[k].ny < 2) return fail("fd2d_batch_units: grid %d size", k);
        const int h = batch_grid_rows(g, grids[k].nx);
        n += (long long)((grids[k].ny + w - 1) / w) * ((grids[k].nx + h - 1) / h);
    }
    if (n > (1LL << 30)) return fail("fd2d_batch_units: too many units");
    *nunits = (int)n;
    return 0;
}

int fd2d_batch_upload(const fd2d_grid *grids, int ngrids, int depth, int rows, int cols,
                      int waves, fd2d_grid *grids_dev, int *units_dev, int max_units,
                      fd2d_batch_desc *out, void *stream)
{
    clear_error();
    if (!grids || ngrids <= 0 || !grids_dev || !units_dev || !out)
        return fail("fd2d_batch_upload: grids, grids_dev, units_dev and out are required");
    BatchGeom g;
    int rc = batch_geom(depth, rows, cols, &g);
    if (rc) return rc;
    bool vec = g.cols == 2;
    int min_wave = INT_MAX, min_probe = INT_MAX;
    for (int k = 0; k < ngrids; ++k) {
        rc = check_grid(grids[k], k);
        if (rc) return rc;
        vec = vec && (grids[k].ny % 2 == 0);
        if (grids[k].src_i >= 0) min_wave = std::min(min_wave, std::max(grids[k].wave_len, 0));
        if (grids[k].probe_i >= 0) min_probe = std::min(min_probe, std::max(grids[k].probe_len, 0));
    }
    const int w = 64 * g.cols - 2 * g.halo;
    // work list: the units whose rows or columns reach a grid edge, a PML
    // band, a source or a probe (they run the general body), then the
    // interior units; dealt out to the waves below
    std::vector<int> edge, inner;
    for (int k = 0; k < ngrids; ++k) {
        const fd2d_grid &G = grids[k];
        const int h = batch_grid_rows(g, G.nx);
        const int ns = (G.ny + w - 1) / w, nb = (G.nx + h - 1) / h;
        const int rlo = G.band_rows[1] > G.band_rows[0] ? std::max(G.band_rows[0], 1) : 0;
        const int rhi = G.band_rows[1] > G.band_rows[0] ? std::min(G.band_rows[1], G.nx - 1) : 0;
        const int clo = G.band_cols[1] > G.band_cols[0] ? std::max(G.band_cols[0], 1) : 0;
        const int chi = G.band_cols[1] > G.band_cols[0] ? std::min(G.band_cols[1], G.ny - 1) : 0;
        for (int by = 0; by < nb; ++by) {
            for (int sx = 0; sx < ns; ++sx) {
                const int js = sx * w - g.halo, i0 = by * h;
                const int i1 = std::min(i0 + h, G.nx);
                const bool brow = !(i0 - 2 * g.depth >= rlo && i1 - 1 + g.depth < rhi);
                const bool bcol = !(js >= clo && js + 64 * g.cols <= chi);
                const bool sp = (G.src_i >= 0 && G.src_i >= i0 - 2 * g.depth && G.src_i <= i1 + g.depth &&
                                 G.src_j >= js && G.src_j < js + 64 * g.cols) ||
                                (G.probe_i >= 0 && G.probe_i >= i0 && G.probe_i < i1 &&
                                 G.probe_j >= js && G.probe_j < js + 64 * g.cols);
                auto &v = (brow || bcol || sp) ? edge : inner;
                v.insert(v.end(), {k, sx, i0, i1});
            }
        }
    }
    const int n = (int)((edge.size() + inner.size()) / 4);
    if (n > max_units) return fail("fd2d_batch_upload: %d units exceed max_units=%d", n, max_units);
    const size_t nedge = edge.size() / 4;
    edge.insert(edge.end(), inner.begin(), inner.end());
    // Balance the waves of the persistent launch: wave `slot` takes the units
    // at slot, slot + nw, slot + 2 nw, ...  The units are sorted by their
    // estimated cost, longest first (rows swept: the block and its 2K rows of
    // warm-up; the general body of edge, PML, source and probe units costs
    // about 1.75 interior rows per row), and dealt out round by round in a
    // snake: round r gives its units to slots 0..nw-1 when r is even and to
    // nw-1..0 when odd, so a wave that drew a long unit in one round draws a
    // short one in the next.  The last, partial round is dealt in order so
    // the list stays dense.  The sort is stable: units of equal cost keep
    // their grid and strip order, so neighbours still run side by side.
    // (MI355X, the 256-grid batch: 330k -> 374k Mcell-updates/s at 128 rows;
    // edge costs 1.0 / 1.4 / 2.5: 360k / 370k / 359k.)
    if (waves <= 0) {
        rc = fd2d_batch_resident(g.depth, g.cols, vec, &waves);
        if (rc) return rc;
    }
    waves = std::max(1, std::min(waves, n));
    {
        const int nw = waves;
        auto cost = [&](int q) {
            const double gen = (size_t)q < nedge ? 1.75 : 1.0;
            return (edge[4 * (size_t)q + 3] - edge[4 * (size_t)q + 2] + 2 * g.depth) * gen;
        };
        std::vector<int> idx(n);
        for (int q = 0; q < n; ++q) idx[q] = q;
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost(a) > cost(b); });
        std::vector<int> lay(4 * (size_t)n);
        const int full = n / nw;
        for (int q = 0; q < n; ++q) {
            const int r = q / nw, i = q % nw;
            const int slot = (r < full && (r & 1)) ? nw - 1 - i : i;
            const size_t pos = (size_t)r * nw + slot;
            for (int c = 0; c < 4; ++c) lay[4 * pos + c] = edge[4 * (size_t)idx[q] + c];
        }
        edge.swap(lay);
    }
    rc = hip_check(hipMemcpyAsync(units_dev, edge.data(), edge.size() * sizeof(int),
                                  hipMemcpyHostToDevice, S(stream)), "units upload");
    if (!rc) rc = hip_check(hipMemcpyAsync(grids_dev, grids, (size_t)ngrids * sizeof(fd2d_grid),
                                           hipMemcpyHostToDevice, S(stream)), "grids upload");
    if (!rc) rc = hip_check(hipStreamSynchronize(S(stream)), "upload sync");
    if (rc) return rc;
    out->ngrids = ngrids;
    out->grids_dev = grids_dev;
    out->units_dev = units_dev;
    out->nunits = n;
    out->depth = g.depth;
    out->rows = g.rows;
    out->cols = g.cols;
    out->waves = waves;
    out->vec = vec ? 1 : 0;
    out->wave_len = min_wave;
    out->probe_len = min_probe;
    out->unit_work = nullptr;
    return 0;
}

int fd2d_batch_run(const fd2d_batch_desc *b, int t0, int nsteps, int max_wave_len,
                   int max_probe_len, void *stream)
{
    clear_error();
    return batch_launches(b, t0, nsteps, max_wave_len, max_probe_len, S(stream));
}

int fd2d_batch_graph(const fd2d_batch_desc *b, int t0, int nsteps, int max_wave_len,
                     int max_probe_len, void *stream, void **graph_exec)
{
    clear_error();
    if (!stream) return fail("fd2d_batch_graph: capture needs a stream other than the null stream");
    if (!graph_exec) return fail("fd2d_batch_graph: graph_exec is required");
    hipStream_t s = S(stream);
    int rc = hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
    if (rc) return rc;
    int lrc = batch_launches(b, t0, nsteps, max_wave_len, max_probe_len, s);
    hipGraph_t graph = nullptr;
    rc = hip_check(hipStreamEndCapture(s, &graph), "end capture");
    if (lrc) {
        if (graph) hipGraphDestroy(graph);
        return lrc;
    }
    if (rc) return rc;
    hipGraphExec_t exec = nullptr;
    rc = hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    hipGraphDestroy(graph);
    if (rc) return rc;
    *graph_exec = exec;
    return 0;
}

int fdtd_graph_launch(void *graph_exec, void *stream)
{
    clear_error();
    if (!graph_exec) return fail("fdtd_graph_launch: graph_exec is NULL");
    return hip_check(hipGraphLaunch(reinterpret_cast<hipGraphExec_t>(graph_exec), S(stream)),
                     "graph launch");
}

int fdtd_graph_destroy(void *graph_exec)
{
    clear_error();
    if (!graph_exec) return 0;
    return hip_check(hipGraphExecDestroy(reinterpret_cast<hipGraphExec_t>(graph_exec)),
                     "graph destroy");
}

}  // extern "C"
