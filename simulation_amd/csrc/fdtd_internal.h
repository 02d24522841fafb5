// fdtd_internal.h -- shared declarations of libfdtd_hip.so (not installed).
//
// Kernel argument structs, the per-thread error message and the launch
// helpers used by the kernel translation units and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/fdtd.h"

namespace fdtd {

// ------------------------------------------------------------ errors
// Records a message for fdtd_last_error() and returns FDTD_EINVAL.
int fail(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
// Checks a HIP status; on failure records "<what>: <hip message>" and
// returns the positive hipError_t.
int hip_check(hipError_t e, const char *what);
// Clears the per-thread message (called at the start of every entry point).
void clear_error();

// ------------------------------------------------------ code-object loading
// The HIP runtime loads a translation unit's device code object at the first
// use of one of its kernels.  Asking for a kernel's attributes is such a use;
// each kernel translation unit exposes one preload_* function that does it
// for a kernel of its own, so fdtd_preload() can load every code object
// before a timed or latency-sensitive run.
inline int preload_kernel(const void *kernel)
{
    hipFuncAttributes a;
    return hip_check(hipFuncGetAttributes(&a, kernel), "fdtd_preload");
}
int preload_fd1d();
int preload_fd2d();
int preload_dform1d();
int preload_tfsf();
int preload_lbc();
int preload_lbc_frame();
int preload_tb();                    // every depth (fd2d_tb.hip)
template <int K>
int preload_tb_depth();

// ------------------------------------------------------ kernel arguments

struct PmlPtrs {
    const float *fx1, *fx2, *fx3;
    const float *fy1, *fy2, *fy3;
    const float *gx2, *gx3;
    const float *gy2, *gy3;
};

struct SrcList {                   // hard sources of one step
    int n;
    int i[FDTD_MAX_SOURCES];
    int j[FDTD_MAX_SOURCES];
    float v[FDTD_MAX_SOURCES];     // values, used when wave is NULL
    const float *wave;             // device [n][wave_len]; value = wave[k*wave_len + t-1]
    int wave_len;
    int t;
};

struct CylList {
    int n;
    double cx[FDTD_MAX_CYLINDERS], cy[FDTD_MAX_CYLINDERS], rad[FDTD_MAX_CYLINDERS];
    double epsr[FDTD_MAX_CYLINDERS], sigma[FDTD_MAX_CYLINDERS];
};

// Arguments of one temporally blocked launch (fd2d_tb.hip).
constexpr int TB_MAX_DEPTH = 12;    // temporally blocked launches: steps per launch
constexpr int MW_MAX_DEPTH = 8;     // multi-wave form: steps per launch
constexpr int LBC_MAX_DEPTH = 10;   // learned boundary: steps per launch (frame tables, LDS)
constexpr int TB_MAX_FREQ = 3;      // fused running DFT: frequencies
constexpr int TB_MAX_LIST_RB = 64;  // boundary launch: whole row blocks listed
constexpr int TB_MAX_LIST_SX = 64;  // boundary launch: whole strips listed
constexpr int TB_MAX_LIST_T = 128;  // boundary launch: single tiles listed

struct TbParams {
    int nx, ny;
    const float *dz_i, *iz_i, *hx_i, *hy_i, *ihx_i, *ihy_i;
    float *dz_o, *iz_o, *hx_o, *hy_o, *ihx_o, *ihy_o, *ez_o;
    const float *naz, *nbz;
    PmlPtrs pml;
    int halo;                      // column halo of the strip geometry (>= depth)
    int rows;                      // rows per block
    // interior band: rows [rlo, rhi) and columns [clo, chi) lie inside
    // [1, n-2] and hold the default profile values (empty when unknown)
    int rlo, rhi, clo, chi;
    // the boundary part's (TB_BOUNDARY) own band: its tiles inside it need
    // neither the edge rows and columns nor profile values (the learned
    // boundary's partition band is wider than that); empty: every tile of
    // the part runs the general bodies
    int erlo, erhi, eclo, echi;
    int t_first;                   // time index of the first fused step
    int nsrc;
    int src_i[FDTD_MAX_SOURCES], src_j[FDTD_MAX_SOURCES];
    const float *src_wave;
    int wave_len;
    int nprobe;
    int probe_i[FDTD_MAX_PROBES], probe_j[FDTD_MAX_PROBES];
    float *probe_out;
    int probe_len;
    int skip_w;                    // leave cells within skip_w of an edge unwritten (0: none)
    // plane wave: TF/SF box rows/columns tf_npml-1 .. n-tf_npml (0: none);
    // step t reads inc[(t - inc_t0 - 1) * inc_stride + k]: k = 0, 1 the
    // incident hx beside the box (hxi[npml-2], hxi[ny-npml]), then the
    // incident ez over the box columns npml-1 .. ny-npml
    int tf_npml, inc_t0, inc_stride;
    const float *inc;
    // free-space tiles (fd2d_tm_desc.tile_work): vac_mode 0 off; 1 the
    // interior tiles record in tile_vac[tile] whether every naz, nbz they
    // read is 1, 0 (as vac_tag | 1, else vac_tag); 2 interior tiles whose
    // entry reads vac_tag | 1 read neither array.  vac_tag changes with every
    // run call (bit 0 clear), so a flag from another call never matches.
    unsigned *tile_vac;
    int vac_mode;
    unsigned vac_tag;
    // running DFT of ez into nf <= TB_MAX_FREQ planes (0: none); tw_c/tw_s
    // [s][f]: twiddles of level s of the launch
    int nf;
    float *ft_r, *ft_i;
    float tw_c[TB_MAX_DEPTH][TB_MAX_FREQ], tw_s[TB_MAX_DEPTH][TB_MAX_FREQ];
    // boundary tiles: the row blocks rb_list (every strip), the strips
    // sx_list (every row block) and single tiles t_list (strip << 16 | row
    // block); together a superset of the tiles that are not interior.  n_rb
    // < 0: no lists (every workgroup takes its tile from the whole grid).
    // A merged launch runs nb boundary workgroups (padded to nb_pad, a
    // multiple of 8), then one per tile of the grid.
    int n_rb, n_sx, n_t, nb, nb_pad;
    int rb_list[TB_MAX_LIST_RB], sx_list[TB_MAX_LIST_SX], t_list[TB_MAX_LIST_T];
};

// Arguments of one learned-boundary frame launch (lbc_frame.hip).
struct LbcFrameParams {
    int nx, ny, K;                              // grid, steps of the launch
    const float *dz_i, *iz_i, *hx_i, *hy_i;     // state at the launch start
    float *dz_o, *iz_o, *hx_o, *hy_o, *ez_o;    // state after K steps (ez_o may be NULL)
    const float *naz, *nbz;
    const float *hist_in;                       // [npos][4][30], newest sample first
    float *hist_out;
    const uint16_t *w;                          // network: bf16 operands
    const float *params;                        //          FP32 parameters (biases, L)
    const void *tab;                            // geometry tables (launch_lbc_frame_setup)
};
int launch_fd2d_energy_diff(int wn, int wm, int nya, int ia0, int ja0, const float *eza,
                            const float *hxa, const float *hya, int nyb, int ib0, int jb0,
                            const float *ezb, const float *hxb, const float *hyb, double *work,
                            hipStream_t s);
// 1-D D/E form (dform1d.hip)
int launch_dxfield(int nx, float *dx, const float *hy, int src_pos, float src_val, int src_on,
                   hipStream_t s);
int launch_exfield_dform(int nx, const fdtd_medium1d &md, const float *dx, float *ix, float *sx,
                         float *ex, hipStream_t s);
int launch_fourier1d(int t, int nf, int nx, float dt, const float *freq, const float *ex, int probe,
                     const fdtd_ftrans &ft, hipStream_t s);
int launch_debye(int nx, float dt, float chi, float tau, float epsr, float sigma,
                 const fdtd_medium1d &md, hipStream_t s);
// 1-D D/E form, fused run (dform1d.hip): DF_W cells per workgroup, up to
// DF_MAXK steps per launch, DFT at up to DF_NF frequencies.
constexpr int DF_NT = 256, DF_W = 512, DF_MAXK = 32, DF_NF = 3;
constexpr int DF_INC_PROBE = 10;    // incident-probe cell of the 1-D DFT (r_in, i_in)
struct DformTbParams {
    int nx, t_first, K;             // steps t_first .. t_first + K - 1
    const float *dx_i, *ix_i, *sx_i, *ex_i, *hy_i;
    float *dx_o, *ix_o, *sx_o, *ex_o, *hy_o;
    const float *bc_i;
    float *bc_o;
    const float *nax, *nbx, *ncx, *ndx;   // ncx NULL: the lossy medium
    int src_pos;                    // < 1: no source
    float src_val[DF_MAXK];
    int nf;
    float *r_pt, *i_pt, *r_in, *i_in;
    float tw_c[DF_MAXK][DF_NF], tw_s[DF_MAXK][DF_NF];
    int nprobe;
    int probe_k[FDTD_MAX_PROBES];
    float *probe_out;
    int probe_len;
};
int launch_dform_tb(const DformTbParams &p, hipStream_t s);
// plane wave (tfsf.hip)
constexpr int INC_CHUNK = 64;       // steps of incident grid per fused chunk
struct IncChunk {
    int n, npml, stride, inc_pos, probe, nf;
    float wave[INC_CHUNK];
    float c[INC_CHUNK][TB_MAX_FREQ], s[INC_CHUNK][TB_MAX_FREQ];
};
int launch_inc_chunk(int ny, float *ezi, float *hxi, float *bc, float *r_in, float *i_in,
                     float *table, const IncChunk &a, hipStream_t s);
void dft_twiddle(int t, float freq, float dt, float *c, float *s);
int launch_ezinct(int ny, float *ezi, const float *hxi, float *bc, int set_inc, int inc_pos,
                  float inc_val, hipStream_t s);
int launch_set(float *p, int pos, float v, hipStream_t s);
int launch_hxinct(int ny, const float *ezi, float *hxi, hipStream_t s);
int launch_inctdz(int nx, int ny, int npml, const float *hxi, float *dz, hipStream_t s);
int launch_incthx(int nx, int ny, int npml, const float *ezi, float *hx, hipStream_t s);
int launch_incthy(int nx, int ny, int npml, const float *ezi, float *hy, hipStream_t s);
int launch_fourier(int t, int nf, int nx, int ny, float dt, const float *freq, const float *ezi,
                   int probe, const float *ez, const fdtd_ftrans &ft, hipStream_t s);

size_t lbc_frame_smem();
// Geometry tables of the frame workgroups (per depth K = 1..8, one after the
// other in the caller's work buffer; built by launch_lbc_frame_setup).
size_t lbc_frame_table_bytes(int nx, int ny, int K);
size_t lbc_frame_table_offset(int nx, int ny, int K);
int launch_lbc_frame_setup(int nx, int ny, int K, void *tab, hipStream_t s);
int lbc_frame_blocks(int nx, int ny, int K);
int launch_lbc_frame(const LbcFrameParams &p, hipStream_t s);

// ----------------------------------------------------- launch wrappers
// (defined in the kernel translation units; return 0 or a hip status)

template <typename T>
int launch_fd1d_exfield(int nx, const T *ca, const T *cb, T *ex, const T *hy,
                        int src_mode, int src_pos, T src_val, hipStream_t s);
template <typename T>
int launch_fd1d_hyfield(int nx, T *ex, T *hy, T *bc, hipStream_t s);
template <typename T>
int launch_fd1d_run(int nx, int nsteps, const T *ca, const T *cb, T *ex, T *hy, T *bc,
                    int src_mode, int src_pos, const T *src_wave,
                    int nprobe, const int *probe_k, T *probe_out, hipStream_t s);
int fd1d_run_lds_limit();          // bytes of LDS the single-workgroup run may use
template <typename T>
int launch_fd1d_dielectric(int nx, double dt, double epsr, double sigma, T *ca, T *cb,
                           hipStream_t s);

int launch_fd2d_dfield(int nx, int ny, const PmlPtrs *pml, float *dz, const float *hx,
                       const float *hy, const SrcList &src, hipStream_t s);
int launch_fd2d_efield(int nx, int ny, const float *naz, const float *nbz,
                       const float *dz, float *iz, float *ez, hipStream_t s);
int launch_fd2d_hfield(int nx, int ny, const PmlPtrs *pml, const float *ez,
                       float *ihx, float *ihy, float *hx, float *hy, hipStream_t s);
int launch_fd2d_probe(int ny, const float *ez, int nprobe, const int *pi, const int *pj,
                      float *out, int stride, int col, hipStream_t s);
int launch_fd2d_pmlparam(int nx, int ny, int npml, const fdtd_pml &pml, hipStream_t s);
int launch_fd2d_dielectric(int nx, int ny, int npml, double dt, const CylList &cyl,
                           int int_sub, float *naz, float *nbz, hipStream_t s);
int launch_fd2d_energy(int nx, int ny, int i0, int i1, int j0, int j1, const float *ez,
                       const float *hx, const float *hy, double *work, hipStream_t s);
// The same reduction with the result written to *out (device).
int launch_fd2d_energy_to(int nx, int ny, int i0, int i1, int j0, int j1, const float *ez,
                          const float *hx, const float *hy, double *work, double *out,
                          hipStream_t s);
// Device address of the library's static energy scratch on the current
// device (FDTD_ENERGY_WORK doubles; looked up once per device).
int energy_scratch(double **out);
size_t energy_scratch_elems();

// Temporally blocked TM launch (fd2d_tb_impl.h): depth in [1,
// tb_max_depth()], cols in {1,2}; part TB_WHOLE: every tile in one launch
// (the boundary tiles first), TB_INTERIOR / TB_BOUNDARY: the interior tiles
// or the others only (two launches that may run concurrently).
enum { TB_INTERIOR = 0, TB_BOUNDARY = 1, TB_WHOLE = 2 };
int tb_max_depth();
bool tb_supported(int depth, int cols);
int launch_fd2d_tb(const TbParams &p, int depth, int cols, bool lossy, bool pml, int part,
                   hipStream_t s);
// one instantiation per depth (fd2d_tb_k<depth>.hip)
template <int K>
int launch_fd2d_tb_depth(const TbParams &p, int cols, bool lossy, bool pml, int part,
                         hipStream_t s);

// Multi-wave form of the whole-grid launch (fd2d_mw_impl.h): interior
// super-tiles swept by MW_WAVES waves together, frame tiles by single waves;
// one column per lane, no plane wave, no DFT.  rows: rows of an interior
// super-tile; frows: rows of a frame tile.
constexpr int MW_WAVES = 4;
int launch_fd2d_mw(const TbParams &p, int depth, bool lossy, bool pml, int rows, int frows,
                   hipStream_t s);
int preload_mw();
// one instantiation per depth (fd2d_mw_k<depth>.hip)
template <int K>
int launch_fd2d_mw_depth(const TbParams &p, bool lossy, bool pml, int rows, int frows,
                         hipStream_t s);
template <int K>
int preload_mw_depth();

// Free-space units of a batched launch: work[u] for unit u, as
// TbParams::tile_vac for tiles (mode 0: none, 1: record, 2: use; entries
// tag | 1 mark units that read only naz = 1, nbz = 0).
struct BatchVac {
    unsigned *work;
    int mode;
    unsigned tag;
};

// Batched grids (lossy PML form): persistent launch over a work list of
// 4-int units {grid, strip, first row, one past the last row}; parity 1
// reads the partner sets.
int launch_fd2d_batch(const fd2d_grid *grids, const int *units, int nunits, int waves, int depth,
                      int cols, bool vec, int parity, int t_first, int write_ez, int halo,
                      BatchVac vac, hipStream_t s);
int fd2d_batch_resident(int depth, int cols, bool vec, int *waves);
template <int K>
int launch_fd2d_batch_depth(const fd2d_grid *grids, const int *units, int nunits, int waves,
                            int cols, bool vec, int parity, int t_first, int write_ez, int halo,
                            BatchVac vac, hipStream_t s);
template <int K>
int fd2d_batch_resident_depth(int cols, bool vec, int *waves);

}  // namespace fdtd
