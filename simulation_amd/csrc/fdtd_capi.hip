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

#include "fdtd_internal.h"

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
};

int plan_tm(const fd2d_tm_desc *d, int nsteps, Plan *pl)
{
    int D = d->depth ? d->depth : 8;
    int C = d->cols ? d->cols : 2;
    if (!tb_supported(D, C)) return fail("fd2d_tm: depth=%d cols=%d not supported", D, C);
    if (64 * C - 2 * D < 8) return fail("fd2d_tm: depth %d too large for %d columns/lane", D, C);
    int H = d->rows;
    if (H == 0) {
        // default from measurements on MI355X (4096^2 .. 16384^2): ~110 row
        // blocks per strip, i.e. several waves per SIMD over the launch
        long h = d->nx / 110;
        if (h < 16) h = 16;
        if (h > 512) h = 512;
        H = (int)h;
    }
    if (H < 1) return fail("fd2d_tm: rows=%d", H);
    // strip halo: the deepest launch's cone, rounded up to whole lanes
    pl->depth = D; pl->cols = C; pl->rows = H; pl->halo = ((D + C - 1) / C) * C;
    pl->launches = 0; pl->tail_unfused = 0;
    if (nsteps <= 0) return 0;
    if (nsteps == 1) { pl->tail_unfused = 1; return 0; }
    int L = (nsteps + D - 1) / D;
    if (L % 2) {
        if (L + 1 <= nsteps) L += 1;
        else { pl->tail_unfused = 1; L = (nsteps - 1 + D - 1) / D; if (L % 2) L += 1; }
    }
    pl->launches = L;
    return 0;
}

// The auxiliary stream on which the boundary part of each temporally blocked
// launch runs, concurrently with the interior part on the caller's stream,
// and the two events that order them.  One set per host thread and device,
// created on first use and kept for the life of the process (no device
// memory is involved).
struct Aux {
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
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
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.fork, hipEventDisableTiming), "aux event");
        if (!rc) rc = hip_check(hipEventCreateWithFlags(&a.join, hipEventDisableTiming), "aux event");
        if (rc) return rc;
    }
    *out = &a;
    return 0;
}

}  // namespace
}  // namespace fdtd

using namespace fdtd;

extern "C" {

int fdtd_abi_version(void) { return FDTD_ABI_VERSION; }

const char *fdtd_last_error(void) { return g_err; }

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
    Aux *ax = nullptr;
    if (L) {
        rc = aux_for_current_device(&ax);
        if (rc) return rc;
    }
    int done = 0;
    for (int l = 0; l < L; ++l) {
        const int k = nfused / L + (l < nfused % L ? 1 : 0);
        float **in = (l % 2 == 0) ? A : B;
        float **out = (l % 2 == 0) ? B : A;
        p.dz_i = in[0]; p.iz_i = in[1]; p.hx_i = in[2]; p.hy_i = in[3]; p.ihx_i = in[4]; p.ihy_i = in[5];
        p.dz_o = out[0]; p.iz_o = out[1]; p.hx_o = out[2]; p.hy_o = out[3]; p.ihx_o = out[4]; p.ihy_o = out[5];
        p.ez_o = (l == L - 1 && !pl.tail_unfused) ? d->ez : nullptr;
        p.t_first = t0 + done + 1;
        // boundary part on the auxiliary stream, after all work queued so far
        // on the caller's stream (which includes the previous interior part)
        rc = hip_check(hipEventRecord(ax->fork, S(stream)), "fork event");
        if (!rc) rc = hip_check(hipStreamWaitEvent(ax->s, ax->fork, 0), "fork wait");
        if (!rc) rc = launch_fd2d_tb(p, k, pl.cols, lossy, pml, 1, ax->s);
        // interior part on the caller's stream, after the previous boundary part
        if (!rc && l > 0) rc = hip_check(hipStreamWaitEvent(S(stream), ax->join, 0), "join wait");
        if (!rc) rc = launch_fd2d_tb(p, k, pl.cols, lossy, pml, 0, S(stream));
        if (!rc) rc = hip_check(hipEventRecord(ax->join, ax->s), "join event");
        if (rc) return rc;
        done += k;
    }
    if (L) {
        rc = hip_check(hipStreamWaitEvent(S(stream), ax->join, 0), "join wait");
        if (rc) return rc;
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

}  // extern "C"
