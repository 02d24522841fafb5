// fd2d_tb_impl.h -- fused, temporally blocked 2-D TM time stepping for gfx950
// (template body; instantiated per depth by fd2d_tb_k<depth>.hip).
//
// One launch advances the whole grid by K leapfrog steps (K = template
// depth).  Semantics per step are those of the reference's per-step kernels
// dfield -> hard source -> efield -> hfield (fd2d/pycuda/fd2d_3_2.py:219-260,
// fd2d/pycuda/fd2d_3_4.py:175-185), operand for operand, so the result equals
// K rounds of fd2d_dfield/efield/hfield bit for bit.
//
// Decomposition (one wavefront per workgroup, no barriers, no LDS):
//   * Columns: a wave owns a strip of S = 64*C columns, lane l holding the C
//     contiguous columns js + l*C .. js + l*C + C-1.  The column left of a
//     lane's first and right of its last come from lanes l-1 / l+1 through
//     DPP wave shifts (one VALU op each).  Each fused step widens the
//     dependency cone by one column on each side, so a strip writes only its
//     central S - 2*halo columns; neighbouring strips overlap by the halo
//     (recomputed, never written twice).
//   * Rows: the wave sweeps down a block of `rows` rows.  Time level s of the
//     K fused steps trails level s-1 by one row: at sweep position p, level s
//     computes the E update of row p-s and the H update of row p-s-1, fed by
//     the H row level s-1 produced in the same iteration and the D row it
//     produced one iteration earlier.  What a level needs from earlier rows is
//     carried in registers, so HBM is read once (level 0, row p) and written
//     once (level K-1) per launch.  A block starts K rows above its first row
//     (warm-up) and ends K rows below its last (drain).
//   * Waves whose rows (columns) all lie inside [1, n-2] and inside the band
//     the caller promises to hold the default PML profile values run a body
//     in which those values are compile-time constants and boundary tests are
//     absent; the few waves at the grid edges or holding a source or probe
//     run the general body.  Multiplying by a constant 1 is exact, so the
//     compiler may drop it without changing any result; the terms 0 * ihx
//     and 0 * ihy of the H update are left out in the interior body, where
//     they cannot change a stored bit (see iterate).
//   * One launch covers a K-step group: the host lists the tiles that are
//     not interior (row blocks, strips, source and probe tiles), their
//     workgroups come first, then one workgroup per tile of the grid in an
//     XCD-aware order, computing its tile if it is interior.  The learned
//     boundary instead runs the interior and the listed tiles as two
//     launches on two streams, beside its frame kernel.
//   * Launches read one state set and write the other (ping-pong), so waves
//     never read a cell another wave has already advanced.
//   * Interior waves of tm_tb_kernel keep two sets of level carries and
//     alternate them between sweep iterations (no register-to-register
//     copies; the batch kernel keeps one set, which holds it at 4 waves per
//     SIMD).  The profile values naz, nbz a level needs are not carried: the
//     sweep keeps those of its last rows in a per-wave LDS ring (NaRing).
//     All waves without DFT planes store through buffer resources
//     whose range check drops the cells a wave does not own, so every
//     iteration issues the same stores and the wait for the next row's loads
//     does not wait for them.
//
// HBM traffic per cell per launch (lossy PML form): reads dz iz hx hy ihx ihy
// naz nbz (32 B, x halo/warm-up overlap), writes dz iz hx hy ihx ihy (24 B,
// + 4 B ez on the last launch of a run), for K cell-updates.
#pragma once

#include <algorithm>

#include "fdtd_internal.h"

namespace fdtd {
namespace tb {

// lane l receives lane l-1's value (lane 0 receives 0)
__device__ __forceinline__ float from_left(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}
// lane l receives lane l+1's value (lane 63 receives 0)
__device__ __forceinline__ float from_right(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                 0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}

template <int C> struct HRow { float hx[C], hy[C], ix[C], iy[C]; };
template <int C> struct DRow { float dz[C], iz[C], na[C], nb[C]; };
// Running-DFT accumulators of a row (NF frequencies; none when NF = 0): they
// follow the D row through the levels, so each DFT plane is read and written
// once per launch.
template <int C, int NF> struct FRow { float r[NF ? NF : 1][C], i[NF ? NF : 1][C]; };

// D state a level hands to the next one.
template <int C> struct DSt { float dz[C], iz[C]; };

// Registers one time level carries from one sweep iteration to the next.
template <int C>
struct Carry {
    HRow<C> h;        // H input of this level at its E row (next iteration: the row above)
    float e[C];       // ez of this level at its E row
    DSt<C> d;         // D state of this level at its E row (input of the next level)
};

// The naz and nbz values of the last K rows the sweep loaded, per lane, in
// LDS: level s reads those of its row p-s there instead of receiving them
// through the carries of levels 0 .. s-1, which would take 2 registers per
// level and two copies per level and row.  Row p goes to slots p % K and
// p % K + K, so level s finds row p-s at slot p % K + K - s: one address per
// sweep position and a constant offset per level.  Each lane reads only what
// it wrote itself.
template <int K, int C>
struct alignas(16) NaRing {
    float v[2 * K][64][2 * C];
    // probe values of the sweep's K levels, written out after the sweep (a
    // global store inside the sweep loop, behind a branch, would turn every
    // iteration's wait for the next row's loads into a wait for all
    // outstanding memory operations)
    float probe[FDTD_MAX_PROBES][K];
};

struct Geom {
    int j0;           // first column of the lane
    int i0, i1;       // owned rows [i0, i1)
    int vj0, vj1;     // owned columns [vj0, vj1)
};

// Element offsets below are 32-bit (the host rejects grids of 2^30 cells or
// more), so every access is a uniform 64-bit array base plus a 32-bit
// per-lane offset.  Loads go through global-address-space pointers: the
// arrays are device memory, and where a base pointer is itself read from
// memory (the batch kernel's grid records) a generic pointer would make the
// loads FLAT, which also count against the LDS counter, so that every wait
// for a loaded row becomes a wait for all outstanding memory operations.
using gfloat = const __attribute__((address_space(1))) float;
typedef float vf2 __attribute__((ext_vector_type(2)));
using gfloat2 = const __attribute__((address_space(1))) vf2;
__device__ __forceinline__ float ldg(const float *p) { return *(gfloat *)p; }
// Element i of a read-only array at an index that is the same in every lane
// (a row profile value): a scalar load through the constant address space,
// so it costs no vector memory instruction and no wait on the vector counter.
using cfloat = const __attribute__((address_space(4))) float;
__device__ __forceinline__ float ldu(const float *base, int i)
{
    return ((cfloat *)base)[__builtin_amdgcn_readfirstlane(i)];
}
__device__ __forceinline__ unsigned ldu(const unsigned *base, int i)
{
    return ((const __attribute__((address_space(4))) unsigned *)base)[__builtin_amdgcn_readfirstlane(i)];
}
__device__ __forceinline__ gfloat *at(const float *base, unsigned off)
{
    const unsigned bytes = off << 2;                  // < 2^32: zero-extended 32-bit offset
    return (gfloat *)(reinterpret_cast<const char *>(base) + bytes);
}
__device__ __forceinline__ float *at(float *base, unsigned off)
{
    const unsigned bytes = off << 2;
    return reinterpret_cast<float *>(reinterpret_cast<char *>(base) + bytes);
}

// Load the lane's C columns of a row (element offset `row` of the row start).
// VEC: one vector load (every lane's columns are wholly inside or outside the
// grid); CLAMP: columns outside the grid read an in-bounds neighbour (their
// values are never used).
template <int C, bool VEC, bool CLAMP>
__device__ __forceinline__ void ld(const float *__restrict__ base, unsigned row, int j0, int ny,
                                   float (&v)[C])
{
    if (VEC) {
        const int jc = CLAMP ? min(max(j0, 0), ny - C) : j0;
        if constexpr (C == 2) {
            const vf2 q = *(gfloat2 *)(at(base, row + (unsigned)jc));
            v[0] = q.x; v[1] = q.y;
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) v[c] = *at(base, row + (unsigned)(jc + c));
        }
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c)
            v[c] = *at(base, row + (unsigned)(CLAMP ? min(max(j0 + c, 0), ny - 1) : j0 + c));
    }
}

template <int C, bool VEC>
__device__ __forceinline__ void st(float *__restrict__ base, unsigned row, int j0, int vj0, int vj1,
                                   const float (&v)[C])
{
    if (VEC) {
        if (j0 >= vj0 && j0 + C <= vj1) {
            if constexpr (C == 2) {
                *reinterpret_cast<float2 *>(at(base, row + (unsigned)j0)) = make_float2(v[0], v[1]);
            } else {
#pragma unroll
                for (int c = 0; c < C; ++c) *at(base, row + (unsigned)(j0 + c)) = v[c];
            }
        }
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c)
            if (j0 + c >= vj0 && j0 + c < vj1) *at(base, row + (unsigned)(j0 + c)) = v[c];
    }
}

// Buffer-resource stores (BS sweeps): the range check drops the lanes and
// rows that are not written (offset OOB) instead of branches, so each sweep
// iteration issues the same number of stores and the wait for the next
// row's loads need not wait for the stores just issued.
constexpr unsigned OOB = 0xFFFFFFFCu;      // >= any grid's byte size (< 2^30 cells)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t out_rsrc(const float *base, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(base), (short)0, (int)bytes,
                                             0x00020000);
}

template <int C>
__device__ __forceinline__ void st_rb(__amdgpu_buffer_rsrc_t rs, unsigned row, int j0, int c0, int c1,
                                      bool okrow, const float (&v)[C])
{
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int j = j0 + c;
        const unsigned off = (okrow && j >= c0 && j < c1) ? (row + (unsigned)j) << 2 : OOB;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v[c]), rs, off, 0, 0);
    }
}

// Loads through a buffer resource: offsets at or past its size read +0
// without touching memory.
template <int C>
__device__ __forceinline__ void ld_rb(__amdgpu_buffer_rsrc_t rs, unsigned row, int j0, float (&v)[C])
{
#pragma unroll
    for (int c = 0; c < C; ++c)
        v[c] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                             rs, (row + (unsigned)(j0 + c)) << 2, 0, 0));
}

// VAC: a free-space tile (every naz 1, nbz 0): neither array is read.
// ZIZ 1: and every iz it reads is +0, so iz is not read either; ZIZ 2: iz
// through buffer resources of P.iz_bytes bytes (0: reads +0, stores
// dropped), so one body serves both.
template <int C, bool LOSSY, bool PML, bool VEC, bool CLAMP, bool VAC = false, int ZIZ = 0,
          class PR>
__device__ __forceinline__ void load_row(const PR &P, int p, int j0, HRow<C> &H, DRow<C> &D)
{
    const int pc = min(max(p, 0), P.nx - 1);     // rows outside the grid: any in-bounds row
    const unsigned row = (unsigned)pc * (unsigned)P.ny;
    ld<C, VEC, CLAMP>(P.dz_i, row, j0, P.ny, D.dz);
    if (VAC) {
#pragma unroll
        for (int c = 0; c < C; ++c) { D.na[c] = 1.f; D.nb[c] = 0.f; }
    } else {
        ld<C, VEC, CLAMP>(P.naz, row, j0, P.ny, D.na);
    }
    ld<C, VEC, CLAMP>(P.hx_i, row, j0, P.ny, H.hx);
    ld<C, VEC, CLAMP>(P.hy_i, row, j0, P.ny, H.hy);
    if (LOSSY) {
        if constexpr (ZIZ == 1) {
#pragma unroll
            for (int c = 0; c < C; ++c) D.iz[c] = 0.f;
        } else if constexpr (ZIZ == 2) {
            ld_rb<C>(out_rsrc(P.iz_i, P.iz_bytes), (unsigned)pc * (unsigned)P.ny, j0, D.iz);
        } else {
            ld<C, VEC, CLAMP>(P.iz_i, row, j0, P.ny, D.iz);
        }
        if (!VAC) ld<C, VEC, CLAMP>(P.nbz, row, j0, P.ny, D.nb);
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c) { D.iz[c] = 0.f; D.nb[c] = 0.f; }
    }
    if (PML) {
        ld<C, VEC, CLAMP>(P.ihx_i, row, j0, P.ny, H.ix);
        ld<C, VEC, CLAMP>(P.ihy_i, row, j0, P.ny, H.iy);
    } else {
#pragma unroll
        for (int c = 0; c < C; ++c) { H.ix[c] = 0.f; H.iy[c] = 0.f; }
    }
}

// Column profile values of the lane (constant along the sweep).  In interior
// strips they are the defaults, as compile-time constants.
template <int C>
struct ColCoef { float gy2[C], gy3[C], fy1[C], fy2[C], fy3[C]; };

template <int C, bool PML, bool BCOL, class PR>
__device__ __forceinline__ void col_coef(const PR &P, int j0, ColCoef<C> &cc)
{
#pragma unroll
    for (int c = 0; c < C; ++c) {
        cc.gy2[c] = 1.f; cc.gy3[c] = 1.f; cc.fy1[c] = 0.f; cc.fy2[c] = 1.f; cc.fy3[c] = 1.f;
        if (PML && BCOL) {
            const int j = min(max(j0 + c, 0), P.ny - 1);
            cc.gy2[c] = ldg(P.pml.gy2 + j); cc.gy3[c] = ldg(P.pml.gy3 + j);
            cc.fy1[c] = ldg(P.pml.fy1 + j); cc.fy2[c] = ldg(P.pml.fy2 + j);
            cc.fy3[c] = ldg(P.pml.fy3 + j);
        }
    }
}

// One sweep iteration at position p: level s computes the E update of row
// p-s and the H update of row p-s-1.  The carries of the previous iteration
// are read from Si (SFi) and this iteration's go to So (SFo); the two may be
// the same set (each level reads its entries before it replaces them).
// BROW: rows may lie outside [1, nx-2] or outside the default band; BCOL: the
// same for the strip's columns; SP: a source or probe lies in the rows and
// columns the wave computes.
// VAC: a free-space tile (naz 1, nbz 0 as constants, no profile ring);
// REC: AND into vac whether the row's naz, nbz, iz are exactly 1, +0, +0;
// ZIZ 1 (with VAC): every iz of the tile is +0 in both state sets, so it is
// neither read nor stored (in free space iz + 0 * ez stays +0); ZIZ 2: iz
// read and stored through buffer resources of P.iz_bytes bytes.
template <int K, int C, bool LOSSY, bool PML, bool VEC, bool BROW, bool BCOL, bool SP, int NF,
          bool BS, bool PRO = false, bool VAC = false, bool REC = false, int ZIZ = 0, class PR>
__device__ __forceinline__ void iterate(const PR &P, const Geom &g, const ColCoef<C> &cc,
                                        NaRing<K, C> &ring,
                                        int p, const HRow<C> &Hload, const DRow<C> &Dload,
                                        const Carry<C> (&Si)[K], Carry<C> (&So)[K],
                                        const FRow<C, NF> &Fload, const FRow<C, NF> (&SFi)[K],
                                        FRow<C, NF> (&SFo)[K], int j, bool &vac)
{
    const int nx = P.nx, ny = P.ny;
    const unsigned lane = threadIdx.x & 63u;
    const unsigned slot = (unsigned)p % (unsigned)K;   // p >= 0
    if constexpr (REC) {
#pragma unroll
        for (int c = 0; c < C; ++c)
            vac = vac && __builtin_bit_cast(unsigned, Dload.na[c]) == 0x3F800000u &&
                  (!LOSSY || (__builtin_bit_cast(unsigned, Dload.nb[c]) == 0u &&
                              __builtin_bit_cast(unsigned, Dload.iz[c]) == 0u));
    }
    if constexpr (!VAC) {
        // row p's profile values (level 0's E row) for the levels below
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float *w = ring.v[slot + h * K][lane];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                w[c] = Dload.na[c];
                w[C + c] = Dload.nb[c];
            }
        }
    }
    HRow<C> Hin = Hload;
    DSt<C> Dpend;                         // this iteration's D output of the previous level
    FRow<C, NF> Fpend;
#pragma unroll
    for (int s = 0; s < K; ++s) {
        const int r = p - s;              // E row of this level
        const int q = r - 1;              // H row
        const int t = P.t_first + s;      // time index of this level
        // D input: the previous level's output of the last iteration, which
        // is then replaced by its output of this iteration
        DRow<C> Din;
        FRow<C, NF> Fin;
        if (s == 0) {
            Din = Dload;
            if constexpr (NF > 0) Fin = Fload;
        } else {
            const float *rd = ring.v[slot + K - s][lane];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                Din.dz[c] = Si[s - 1].d.dz[c];
                Din.iz[c] = ZIZ == 1 ? 0.f : Si[s - 1].d.iz[c];
                Din.na[c] = VAC ? 1.f : rd[c];
                Din.nb[c] = (LOSSY && !VAC) ? rd[C + c] : 0.f;
            }
            So[s - 1].d = Dpend;
            if constexpr (NF > 0) {
                Fin = SFi[s - 1];
                SFo[s - 1] = Fpend;
            }
        }
        if (PRO && j < 2 * s + 1) {
            // warm-up position j of the sweep: level s contributes to the
            // tile's outputs only from position 2s+1 on (its E row p-s lies
            // above the dependency cone of the first output row), so it only
            // passes the H row it would carry; the other carries it leaves
            // are not read before it has written them
            So[s].h = Hin;
            continue;
        }
        const HRow<C> Hp = Si[s].h;
        float gx2r = 1.f, gx3r = 1.f, fx1q = 0.f, fx2q = 1.f, fx3q = 1.f;
        if (PML && BROW) {
            const int rc = min(max(r, 0), nx - 1), qc = min(max(q, 0), nx - 1);
            gx2r = ldu(P.pml.gx2, rc); gx3r = ldu(P.pml.gx3, rc);
            fx1q = ldu(P.pml.fx1, qc); fx2q = ldu(P.pml.fx2, qc); fx3q = ldu(P.pml.fx3, qc);
        }
        // ------------------------------------------------ E update, row r
        const float hxl = from_left(Hin.hx[C - 1]);
        DSt<C> Dn;
        float En[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float hL = (c == 0) ? hxl : Hin.hx[c - 1];
            const float curl = Hin.hy[c] - Hp.hy[c] - Hin.hx[c] + hL;
            float dz;
            if (PML) dz = gx3r * cc.gy3[c] * Din.dz[c] + gx2r * cc.gy2[c] * 0.5f * curl;
            else dz = Din.dz[c] + 0.5f * curl;
            if (BCOL) dz = (g.j0 + c >= 1) ? dz : Din.dz[c];     // column 0: no update
            if (BROW) dz = (r >= 1) ? dz : Din.dz[c];            // row 0: no update
            Dn.dz[c] = dz;
        }
        for (int k = 0; SP && k < P.nsrc; ++k) {
            if (r == P.src_i[k]) {
                // a scalar load (the same address in every lane): no vector
                // load behind a branch in the loop, so the waits for the
                // prefetched row stay counted
                const float v = ldu(P.src_wave + (size_t)k * P.wave_len, t - 1);
#pragma unroll
                for (int c = 0; c < C; ++c)
                    if (g.j0 + c == P.src_j[k]) Dn.dz[c] = v;
            }
        }
        // plane wave: the TF/SF correction of dz on the box's columns
        // (fd2d/pycuda/fd2d_3_4.py:164-170), after the D update
        if (SP && NF > 0 && P.tf_npml && r >= P.tf_npml - 1 && r <= nx - P.tf_npml) {
            const float *iv = P.inc + (size_t)(t - P.inc_t0 - 1) * P.inc_stride;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const int j = g.j0 + c;
                if (j == P.tf_npml - 1) Dn.dz[c] = Dn.dz[c] + 0.5f * iv[0];
                if (j == ny - P.tf_npml) Dn.dz[c] = Dn.dz[c] - 0.5f * iv[1];
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float dz = Dn.dz[c];
            float iz = Din.iz[c], ez;
            if (LOSSY) {
                ez = Din.na[c] * (dz - iz);
                iz = iz + Din.nb[c] * ez;
            } else {
                ez = Din.na[c] * dz;
            }
            Dn.iz[c] = iz;
            En[c] = ez;
        }
        FRow<C, NF> Fn;
        if constexpr (NF > 0) {
#pragma unroll
            for (int f = 0; f < NF; ++f)
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    Fn.r[f][c] = Fin.r[f][c] + P.tw_c[s][f] * En[c];
                    Fn.i[f][c] = Fin.i[f][c] - P.tw_s[s][f] * En[c];
                }
        }
        if (SP && P.nprobe && r >= g.i0 && r < g.i1) {
            for (int pr = 0; pr < P.nprobe; ++pr)
                if (r == P.probe_i[pr]) {
#pragma unroll
                    for (int c = 0; c < C; ++c) {
                        const int j = g.j0 + c;
                        if (j == P.probe_j[pr] && j >= g.vj0 && j < g.vj1) ring.probe[pr][s] = En[c];
                    }
                }
        }
        // ------------------------------------------------ H update, row q
        HRow<C> Hout;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            const float e = Si[s].e[c];
            // the column right of the lane's last: the next lane's first,
            // shifted across when it is used rather than carried
            const float eR = (c + 1 < C) ? Si[s].e[c + 1] : from_right(Si[s].e[0]);
            const float eD = En[c];
            float hx, hy, ax, ay;
            if (PML && !BROW && !BCOL) {
                // 0.5 * eR is the right column's 0.5 * e, shifted across as
                // it is (the same bits): the shift then folds into the
                // subtraction, as the one of eR does into e - eR
                const float he = 0.5f * e;
                const float heR = (c + 1 < C) ? 0.5f * Si[s].e[c + 1] : from_right(0.5f * Si[s].e[0]);
                // every profile value of the tile is its default (fx1 = fy1
                // = 0, the others 1): the terms 0 * ihx and 0 * ihy only add
                // a zero.  x + (+-0) is x for every x but a zero, and only
                // the sign of a zero result could differ: an interior hx or
                // hy is never -0 (they start at +0, and hx + t or hy - t is
                // -0 only from -0), so the stored bits are the same as with
                // the terms (the sum differs from the reference only when
                // ihx or ihy is infinite or NaN, in a run that has already
                // blown up).
                ax = Hp.ix[c] + (e - eR);
                ay = Hp.iy[c] + (e - eD);
                hx = Hp.hx[c] + (he - heR);
                hy = Hp.hy[c] - (he - 0.5f * eD);
            } else if (PML) {
                ax = Hp.ix[c] + (e - eR);
                ay = Hp.iy[c] + (e - eD);
                hx = cc.fy3[c] * Hp.hx[c] + cc.fy2[c] * (0.5f * e - 0.5f * eR + fx1q * ax);
                hy = fx3q * Hp.hy[c] - fx2q * (0.5f * e - 0.5f * eD + cc.fy1[c] * ay);
            } else {
                ax = Hp.ix[c];
                ay = Hp.iy[c];
                hx = Hp.hx[c] + 0.5f * (e - eR);
                hy = Hp.hy[c] - 0.5f * (e - eD);
            }
            if (BCOL || BROW) {
                bool upd = true;
                if (BCOL) upd = (g.j0 + c <= ny - 2);           // column ny-1: no update
                if (BROW) upd = upd && (q <= nx - 2);            // row nx-1: no update
                hx = upd ? hx : Hp.hx[c];
                hy = upd ? hy : Hp.hy[c];
                ax = upd ? ax : Hp.ix[c];
                ay = upd ? ay : Hp.iy[c];
            }
            Hout.hx[c] = hx; Hout.hy[c] = hy; Hout.ix[c] = ax; Hout.iy[c] = ay;
        }
        // plane wave: TF/SF corrections of hx on the box's side columns and
        // of hy on its top and bottom rows (fd2d/pycuda/fd2d_3_4.py:211-226)
        if (SP && NF > 0 && P.tf_npml) {
            const int a = P.tf_npml - 1, b = ny - P.tf_npml, qb = nx - P.tf_npml;
            const float *iv = P.inc + (size_t)(t - P.inc_t0 - 1) * P.inc_stride + 2 - a;
            if (q >= a && q <= qb) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int j = g.j0 + c;
                    if (j == P.tf_npml - 2) Hout.hx[c] = Hout.hx[c] + 0.5f * iv[a];
                    if (j == b) Hout.hx[c] = Hout.hx[c] - 0.5f * iv[b];
                }
            }
            if (q == P.tf_npml - 2 || q == qb) {
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const int j = g.j0 + c;
                    if (j >= a && j <= b) {
                        if (q == P.tf_npml - 2) Hout.hy[c] = Hout.hy[c] - 0.5f * iv[j];
                        if (q == qb) Hout.hy[c] = Hout.hy[c] + 0.5f * iv[j];
                    }
                }
            }
        }
        // carry this level's rows to the next iteration
        So[s].h = Hin;
#pragma unroll
        for (int c = 0; c < C; ++c) So[s].e[c] = En[c];
        if (s + 1 < K) {
            Hin = Hout;          // input of the next level
#pragma unroll
            for (int c = 0; c < C; ++c) {
                Dpend.dz[c] = Dn.dz[c];
                Dpend.iz[c] = Dn.iz[c];
            }
            if constexpr (NF > 0) Fpend = Fn;
        } else {
            // last level: D/E of row r and H of row q are final for this launch
            // (cells within skip_w of an edge are left to another kernel)
            const int sw = P.skip_w;
            const int c0 = max(g.vj0, sw), c1 = min(g.vj1, ny - sw);
            if constexpr (BS) {
                static_assert(NF == 0, "buffer stores: no DFT planes");
                const unsigned nb = (unsigned)P.nx * (unsigned)ny * 4u;   // < 2^32
                const bool okr = r >= g.i0 && r < g.i1 && r >= sw && r < P.nx - sw;
                const unsigned rowr = (unsigned)r * (unsigned)ny;
                st_rb<C>(out_rsrc(P.dz_o, nb), rowr, g.j0, c0, c1, okr, Dn.dz);
                if constexpr (LOSSY && ZIZ == 2)
                    st_rb<C>(out_rsrc(P.iz_o, P.iz_bytes), rowr, g.j0, c0, c1, okr, Dn.iz);
                else if constexpr (LOSSY && ZIZ == 0)
                    st_rb<C>(out_rsrc(P.iz_o, nb), rowr, g.j0, c0, c1, okr, Dn.iz);
                st_rb<C>(out_rsrc(P.ez_o, P.ez_o ? nb : 0u), rowr, g.j0, c0, c1, okr, En);
                const bool okq = q >= g.i0 && q < g.i1 && q >= sw && q < P.nx - sw;
                const unsigned rowq = (unsigned)q * (unsigned)ny;
                st_rb<C>(out_rsrc(P.hx_o, nb), rowq, g.j0, c0, c1, okq, Hout.hx);
                st_rb<C>(out_rsrc(P.hy_o, nb), rowq, g.j0, c0, c1, okq, Hout.hy);
                if (PML) {
                    st_rb<C>(out_rsrc(P.ihx_o, nb), rowq, g.j0, c0, c1, okq, Hout.ix);
                    st_rb<C>(out_rsrc(P.ihy_o, nb), rowq, g.j0, c0, c1, okq, Hout.iy);
                }
            } else {
                if (r >= g.i0 && r < g.i1 && r >= sw && r < P.nx - sw) {
                    const unsigned row = (unsigned)r * (unsigned)ny;
                    st<C, VEC>(P.dz_o, row, g.j0, c0, c1, Dn.dz);
                    static_assert(ZIZ != 2, "iz through buffer resources: buffer-store sweeps only");
                    if (LOSSY && ZIZ == 0) st<C, VEC>(P.iz_o, row, g.j0, c0, c1, Dn.iz);
                    if (P.ez_o) st<C, VEC>(P.ez_o, row, g.j0, c0, c1, En);
                    if constexpr (NF > 0) {
                        const size_t plane = (size_t)P.nx * (size_t)ny;
#pragma unroll
                        for (int f = 0; f < NF; ++f)
                            if (f < P.nf) {
                                st<C, VEC>(P.ft_r + f * plane, row, g.j0, c0, c1, Fn.r[f]);
                                st<C, VEC>(P.ft_i + f * plane, row, g.j0, c0, c1, Fn.i[f]);
                            }
                    }
                }
                if (q >= g.i0 && q < g.i1 && q >= sw && q < P.nx - sw) {
                    const unsigned row = (unsigned)q * (unsigned)ny;
                    st<C, VEC>(P.hx_o, row, g.j0, c0, c1, Hout.hx);
                    st<C, VEC>(P.hy_o, row, g.j0, c0, c1, Hout.hy);
                    if (PML) {
                        st<C, VEC>(P.ihx_o, row, g.j0, c0, c1, Hout.ix);
                        st<C, VEC>(P.ihy_o, row, g.j0, c0, c1, Hout.iy);
                    }
                }
            }
        }
    }
}

// DFT accumulators of row p (planes f < P.nf; rows outside the grid read an
// in-bounds row, never stored)
template <int C, int NF, bool VEC, bool CLAMP, class PR>
__device__ __forceinline__ void load_ft(const PR &P, int p, int j0, FRow<C, NF> &F)
{
    const int pc = min(max(p, 0), P.nx - 1);
    const unsigned row = (unsigned)pc * (unsigned)P.ny;
    const size_t plane = (size_t)P.nx * (size_t)P.ny;
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        if (f < P.nf) {
            ld<C, VEC, CLAMP>(P.ft_r + f * plane, row, j0, P.ny, F.r[f]);
            ld<C, VEC, CLAMP>(P.ft_i + f * plane, row, j0, P.ny, F.i[f]);
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) { F.r[f][c] = 0.f; F.i[f][c] = 0.f; }
        }
    }
}

// Returns, with REC, whether every naz, nbz, iz the sweep read is exactly 1,
// +0, +0 (wave-uniform); otherwise false.
template <int K, int C, bool LOSSY, bool PML, bool VEC, bool BROW, bool BCOL, bool SP, int NF = 0,
          bool BS = false, bool PP = false, bool VAC = false, bool REC = false, int ZIZ = 0,
          class PR>
__device__ __forceinline__ bool sweep(const PR &P, const Geom &g, NaRing<K, C> &ring)
{
    bool vac = true;
    ColCoef<C> cc;
    col_coef<C, PML, BCOL>(P, g.j0, cc);
    if constexpr (!VAC) {
        // rows above the sweep's first read zeros, as the carries start
        const unsigned lane = threadIdx.x & 63u;
#pragma unroll
        for (int k = 0; k < 2 * K; ++k)
#pragma unroll
            for (int c = 0; c < 2 * C; ++c) ring.v[k][lane][c] = 0.f;
    }
    // PP: two carry sets, St[0] and St[1]; an iteration reads one and writes
    // the other, so no level's carry is copied from register to register
    // (fewer instructions, more registers).  Otherwise one set, updated in
    // place.
    constexpr int NS = PP ? 2 : 1;
    Carry<C> St[NS][K];
    FRow<C, NF> SF[NS][K];
#pragma unroll
    for (int a = 0; a < NS; ++a)
#pragma unroll
        for (int s = 0; s < K; ++s) {
#pragma unroll
            for (int c = 0; c < C; ++c) {
                St[a][s].h.hx[c] = 0.f; St[a][s].h.hy[c] = 0.f;
                St[a][s].h.ix[c] = 0.f; St[a][s].h.iy[c] = 0.f;
                St[a][s].e[c] = 0.f;
                St[a][s].d.dz[c] = 0.f; St[a][s].d.iz[c] = 0.f;
            }
            if constexpr (NF > 0) {
#pragma unroll
                for (int f = 0; f < NF; ++f)
#pragma unroll
                    for (int c = 0; c < C; ++c) { SF[a][s].r[f][c] = 0.f; SF[a][s].i[f][c] = 0.f; }
            }
        }
    const int pbeg = max(g.i0 - K, 0);
    const int pend = g.i1 - 1 + K;
    // row p+1 is in flight while row p is computed (rows past the grid read a
    // clamped row)
    HRow<C> H0, H1;
    DRow<C> D0, D1;
    FRow<C, NF> F0, F1;
    load_row<C, LOSSY, PML, VEC, BCOL, VAC, ZIZ>(P, pbeg, g.j0, H0, D0);
    if constexpr (NF > 0) load_ft<C, NF, VEC, BCOL>(P, pbeg, g.j0, F0);
    // BS: the first row is needed at once; waiting for it outside the loop
    // keeps the loop head free of a wait that would also catch the stores of
    // the previous iteration
    if constexpr (BS) __builtin_amdgcn_s_waitcnt(0);
    if constexpr (PP) {
        static_assert(!PP || (!BROW && NF == 0), "two carry sets: interior tiles only");
        int p = pbeg;
        // interior tiles start K rows above their first row (pbeg = i0-K);
        // in the first 2K-1 positions level s only starts at position
        // i0-K+2s+1 (one carry set; the sweep then continues in St[0])
        const int c0 = g.i0 - K;
#pragma unroll 1
        for (; p < c0 + 2 * K - 1; ++p) {
            load_row<C, LOSSY, PML, VEC, BCOL, VAC, ZIZ>(P, p + 1, g.j0, H1, D1);
            iterate<K, C, LOSSY, PML, VEC, BROW, BCOL, SP, NF, BS, true, VAC, REC, ZIZ>(
                P, g, cc, ring, p, H0, D0, St[0], St[0], F0, SF[0], SF[0], p - c0, vac);
            H0 = H1;
            D0 = D1;
        }
        // two rows per trip, so the row registers alternate as well
        for (; p < pend; p += 2) {
            load_row<C, LOSSY, PML, VEC, BCOL, VAC, ZIZ>(P, p + 1, g.j0, H1, D1);
            if constexpr (NF > 0) load_ft<C, NF, VEC, BCOL>(P, p + 1, g.j0, F1);
            iterate<K, C, LOSSY, PML, VEC, BROW, BCOL, SP, NF, BS, false, VAC, REC, ZIZ>(
                P, g, cc, ring, p, H0, D0, St[0], St[NS - 1], F0, SF[0], SF[NS - 1], 0, vac);
            load_row<C, LOSSY, PML, VEC, BCOL, VAC, ZIZ>(P, p + 2, g.j0, H0, D0);
            if constexpr (NF > 0) load_ft<C, NF, VEC, BCOL>(P, p + 2, g.j0, F0);
            iterate<K, C, LOSSY, PML, VEC, BROW, BCOL, SP, NF, BS, false, VAC, REC, ZIZ>(
                P, g, cc, ring, p + 1, H1, D1, St[NS - 1], St[0], F1, SF[NS - 1], SF[0], 0, vac);
        }
        if (p == pend)
            iterate<K, C, LOSSY, PML, VEC, BROW, BCOL, SP, NF, BS, false, VAC, REC, ZIZ>(
                P, g, cc, ring, p, H0, D0, St[0], St[NS - 1], F0, SF[0], SF[NS - 1], 0, vac);
    } else {
        for (int p = pbeg; p <= pend; ++p) {
            load_row<C, LOSSY, PML, VEC, BCOL, VAC, ZIZ>(P, p + 1, g.j0, H1, D1);
            if constexpr (NF > 0) load_ft<C, NF, VEC, BCOL>(P, p + 1, g.j0, F1);
            iterate<K, C, LOSSY, PML, VEC, BROW, BCOL, SP, NF, BS, false, VAC, REC, ZIZ>(
                P, g, cc, ring, p, H0, D0, St[0], St[0], F0, SF[0], SF[0], 0, vac);
            H0 = H1;
            D0 = D1;
            if constexpr (NF > 0) F0 = F1;
        }
    }
    if constexpr (SP) {
        // the recorded probe values: probe pr's K values if its cell is one
        // this wave owns (row in [i0, i1), column in [vj0, vj1)); lane s
        // writes level s.  The fences make the owning lane's LDS writes
        // visible to the other lanes of the wave.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int lane = (int)(threadIdx.x & 63u);
        const int js = g.j0 - lane * C;           // first column of the strip
        for (int pr = 0; pr < P.nprobe; ++pr) {
            const int pi = P.probe_i[pr], pj = P.probe_j[pr];
            if (pi >= g.i0 && pi < g.i1 && pj >= g.vj0 && pj < g.vj1 && pj >= js &&
                pj < js + 64 * C && lane < K)
                *(__attribute__((address_space(1))) float *)(P.probe_out + (size_t)pr * P.probe_len +
                                                             (P.t_first + lane - 1)) =
                    ring.probe[pr][lane];
        }
    }
    if constexpr (REC) return __builtin_amdgcn_ballot_w64(!vac) == 0;
    return false;
}

// PART TB_WHOLE: one launch over every tile -- the listed boundary tiles
// first (grid edges, PML bands, sources and probes: the general bodies,
// slower, so they start early), then one workgroup per tile of the grid,
// which computes its tile if it is interior (any tile when there are no
// lists).  TB_INTERIOR / TB_BOUNDARY: the interior tiles (whole grid) or the
// listed others, as two launches that write disjoint cells and may run
// concurrently.  A wave returns at once if its tile is not its part's.
template <int K, int C, bool LOSSY, bool PML, bool VEC, int PART, int NF = 0>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu(C == 1 && K <= 11 && NF == 0 ? 3 : 1)))
tm_tb_kernel(const TbParams P)
{
    constexpr int S = 64 * C;
    const int w = S - 2 * P.halo;
    const int nsx = (P.ny + w - 1) / w, nrb = (P.nx + P.rows - 1) / P.rows;
    // which tiles this workgroup may compute, and where its index comes from
    bool want_int = PART != TB_BOUNDARY, want_bnd = PART != TB_INTERIOR;
    int bi = -1;                     // index into the boundary lists
    unsigned L = blockIdx.x, nt = gridDim.x;
    if (PART == TB_BOUNDARY) {
        if (P.n_rb >= 0) bi = (int)blockIdx.x;
    } else if (PART == TB_WHOLE) {
        if ((int)blockIdx.x < P.nb_pad) {
            if ((int)blockIdx.x >= P.nb) return;
            bi = (int)blockIdx.x;
            want_int = false;
        } else {
            L = blockIdx.x - (unsigned)P.nb_pad;       // nb_pad % 8 == 0: the XCD of L
            nt = gridDim.x - (unsigned)P.nb_pad;
            want_bnd = P.n_rb < 0;                     // no lists: every tile here
        }
    }
    int bx, by;
    if (bi >= 0) {
        // listed row blocks, then listed strips, then single tiles; a tile
        // listed twice is left to its first entry
        int i = bi;
        auto in_rb = [&](int b) {
            for (int k = 0; k < P.n_rb; ++k)
                if (P.rb_list[k] == b) return true;
            return false;
        };
        auto in_sx = [&](int x) {
            for (int k = 0; k < P.n_sx; ++k)
                if (P.sx_list[k] == x) return true;
            return false;
        };
        if (i < P.n_rb * nsx) {
            by = P.rb_list[i / nsx];
            bx = i % nsx;
        } else if ((i -= P.n_rb * nsx) < P.n_sx * nrb) {
            bx = P.sx_list[i / nrb];
            by = i % nrb;
            if (in_rb(by)) return;
        } else {
            i -= P.n_sx * nrb;
            bx = P.t_list[i] >> 16;
            by = P.t_list[i] & 0xffff;
            if (in_rb(by) || in_sx(bx)) return;
        }
    } else {
        // XCD-aware tile order: the dispatcher hands workgroup L to XCD
        // L % 8.  The workgroups of one XCD take a contiguous run of tiles,
        // strips fastest, so neighbouring strips -- which read each other's
        // halo columns at the same rows at about the same time -- and the
        // row blocks below them share that XCD's L2.
        const unsigned xcd = L % 8u, q = nt / 8u, rem = nt % 8u;
        const unsigned tile = xcd * q + min(xcd, rem) + L / 8u;
        bx = (int)(tile % (unsigned)nsx);
        by = (int)(tile / (unsigned)nsx);
    }
    const int js = bx * (S - 2 * P.halo) - P.halo;   // first column of the strip
    Geom g;
    g.j0 = js + (int)threadIdx.x * C;
    g.i0 = by * P.rows;
    g.i1 = min(g.i0 + P.rows, P.nx);
    g.vj0 = js + P.halo;
    g.vj1 = min(js + S - P.halo, P.ny);
    // every row the sweep touches, [i0-2K, i1-1+K], inside the interior band?
    const bool brow = !(g.i0 - 2 * K >= P.rlo && g.i1 - 1 + K < P.rhi);
    // every column of the strip inside the interior band?
    const bool bcol = !(js >= P.clo && js + S <= P.chi);
    // a source or probe inside the rows and columns this wave computes?
    bool sp = false;
    for (int k = 0; k < P.nsrc; ++k)
        sp |= P.src_i[k] >= g.i0 - 2 * K && P.src_i[k] <= g.i1 + K && P.src_j[k] >= js &&
              P.src_j[k] < js + S;
    for (int k = 0; k < P.nprobe; ++k)
        sp |= P.probe_i[k] >= g.i0 && P.probe_i[k] < g.i1 && P.probe_j[k] >= js &&
              P.probe_j[k] < js + S;
    if (P.tf_npml) {
        // a TF/SF-corrected cell in the rows and columns this wave computes:
        // rows npml-2 .. nx-npml on columns npml-2, npml-1, ny-npml, or rows
        // npml-2 and nx-npml on columns npml-1 .. ny-npml
        const int m = P.tf_npml, rlo = g.i0 - 2 * K, rhi = g.i1 + K;
        auto row_in = [&](int r) { return r >= rlo && r <= rhi; };
        auto col_in = [&](int j) { return j >= js && j < js + S; };
        sp |= rhi >= m - 2 && rlo <= P.nx - m && (col_in(m - 2) || col_in(m - 1) || col_in(P.ny - m));
        sp |= (row_in(m - 2) || row_in(P.nx - m)) && js + S > m - 1 && js <= P.ny - m;
    }
    const bool interior = !brow && !bcol && !sp;
    __shared__ NaRing<K, C> ring;
    // free-space tiles (tile index tid): recorded by a run call's first
    // launch, read by the later ones.  The recording launch writes every
    // tile's entry -- a tile that is not interior (at its depth) records "not
    // free space" -- so a later launch of smaller depth, which may find more
    // tiles interior, never reads an entry this recording did not write.
    const unsigned tid = (unsigned)by * (unsigned)nsx + (unsigned)bx;
    if constexpr (NF == 0 && PART != TB_BOUNDARY) {
        if (P.vac_mode == 1 && !interior && bi < 0 && threadIdx.x == 0) P.tile_vac[tid] = P.vac_tag;
    }
    if (interior) {
        if (want_int) {
            if constexpr (NF == 0 && PART != TB_BOUNDARY) {
                if (P.vac_mode == 2 && ldu(P.tile_vac, (int)tid) == (P.vac_tag | 1u)) {
                    // iz is +0 there in both state sets: the recording launch
                    // read +0 and wrote +0, and iz + 0 * ez stays +0
                    sweep<K, C, LOSSY, PML, VEC, false, false, false, NF, true, C == 1, true, false,
                          LOSSY ? 1 : 0>(P, g, ring);
                } else {
                    const bool v =
                        sweep<K, C, LOSSY, PML, VEC, false, false, false, NF, true, C == 1, false, true>(
                            P, g, ring);
                    if (P.vac_mode == 1 && threadIdx.x == 0)
                        P.tile_vac[tid] = P.vac_tag | (v ? 1u : 0u);
                }
            } else {
                sweep<K, C, LOSSY, PML, VEC, false, false, false, NF, NF == 0, NF == 0 && C == 1>(P, g, ring);
            }
        }
    } else if (want_bnd) {
        if constexpr (PART == TB_BOUNDARY && NF == 0) {
            // a tile of the boundary part inside the part's own band runs
            // the interior body (the learned boundary's part 1: tiles near
            // the frame but clear of the grid's edge rows and columns)
            const bool ebrow = !(g.i0 - 2 * K >= P.erlo && g.i1 - 1 + K < P.erhi);
            const bool ebcol = !(js >= P.eclo && js + S <= P.echi);
            if (!ebrow && !ebcol && !sp) {
                sweep<K, C, LOSSY, PML, VEC, false, false, false, NF, true, C == 1>(P, g, ring);
                return;
            }
        }
        if (!brow && !sp) sweep<K, C, LOSSY, PML, VEC, false, true, false, NF, NF == 0>(P, g, ring);
        else sweep<K, C, LOSSY, PML, VEC, true, true, true, NF, NF == 0>(P, g, ring);
    }
}

// The tiles of the boundary launch as lists (TbParams::rb_list, sx_list,
// t_list), from the same predicates as tm_tb_kernel: row blocks with brow or
// a TF/SF row in reach, strips with bcol or a TF/SF column, and the tiles of
// sources and probes elsewhere.  Returns the number of workgroups of the
// compact launch, or -1 if a list would overflow (q.n_rb = -1: whole grid).
template <int K, int C>
int tb_tile_lists(TbParams &q)
{
    constexpr int S = 64 * C;
    const int w = S - 2 * q.halo;
    const int nsx = (q.ny + w - 1) / w, nrb = (q.nx + q.rows - 1) / q.rows;
    q.n_rb = q.n_sx = q.n_t = 0;
    auto fail_list = [&]() { q.n_rb = -1; q.n_sx = q.n_t = 0; return -1; };
    if (nsx >= 0x10000 || nrb >= 0x10000) return fail_list();    // t_list packs 16 + 16 bits
    const int m = q.tf_npml;
    for (int b = 0; b < nrb; ++b) {
        const int i0 = b * q.rows, i1 = std::min(i0 + q.rows, q.nx);
        bool in = !(i0 - 2 * K >= q.rlo && i1 - 1 + K < q.rhi);
        if (m) {
            const int lo = i0 - 2 * K, hi = i1 + K;
            in = in || (m - 2 >= lo && m - 2 <= hi) || (q.nx - m >= lo && q.nx - m <= hi);
        }
        if (in) {
            if (q.n_rb == TB_MAX_LIST_RB) return fail_list();
            q.rb_list[q.n_rb++] = b;
        }
    }
    for (int x = 0; x < nsx; ++x) {
        const int js = x * w - q.halo;
        auto col_in = [&](int j) { return j >= js && j < js + S; };
        bool in = !(js >= q.clo && js + S <= q.chi);
        if (m) in = in || col_in(m - 2) || col_in(m - 1) || col_in(q.ny - m);
        if (in) {
            if (q.n_sx == TB_MAX_LIST_SX) return fail_list();
            q.sx_list[q.n_sx++] = x;
        }
    }
    auto listed = [&](int x, int b) {
        for (int k = 0; k < q.n_rb; ++k)
            if (q.rb_list[k] == b) return true;
        for (int k = 0; k < q.n_sx; ++k)
            if (q.sx_list[k] == x) return true;
        for (int k = 0; k < q.n_t; ++k)
            if (q.t_list[k] == (x << 16 | b)) return true;
        return false;
    };
    // cells (i, j) whose tiles are single entries: a source reaches rows
    // [i0 - 2K, i1 + K] of a tile, a probe only its own rows
    auto add_cell = [&](int ci, int cj, int below, int above) {
        for (int b = 0; b < nrb; ++b) {
            const int i0 = b * q.rows, i1 = std::min(i0 + q.rows, q.nx);
            if (!(ci >= i0 - below && ci < i1 + above)) continue;
            for (int x = 0; x < nsx; ++x) {
                const int js = x * w - q.halo;
                if (!(cj >= js && cj < js + S) || listed(x, b)) continue;
                if (q.n_t == TB_MAX_LIST_T) return false;
                q.t_list[q.n_t++] = x << 16 | b;
            }
        }
        return true;
    };
    for (int k = 0; k < q.nsrc; ++k)
        if (!add_cell(q.src_i[k], q.src_j[k], 2 * K, K + 1)) return fail_list();
    for (int k = 0; k < q.nprobe; ++k)
        if (!add_cell(q.probe_i[k], q.probe_j[k], 0, 0)) return fail_list();
    return q.n_rb * nsx + q.n_sx * nrb + q.n_t;
}

template <int K, int C, bool VEC, int PART>
int launch_kv(const TbParams &p_in, bool lossy, bool pml, hipStream_t s)
{
    constexpr int S = 64 * C;
    const int w = S - 2 * p_in.halo;
    const long ntile = (long)((p_in.ny + w - 1) / w) * (long)((p_in.nx + p_in.rows - 1) / p_in.rows);
    TbParams p = p_in;
    p.n_rb = -1;
    p.n_sx = p.n_t = 0;
    p.nb = p.nb_pad = 0;
    long nwg = ntile;
    if (PART != TB_INTERIOR) {
        const int n = tb_tile_lists<K, C>(p);
        if (PART == TB_BOUNDARY) {
            if (n == 0) return 0;             // every tile is interior
            if (n > 0) nwg = n;
        } else if (n > 0) {
            p.nb = n;
            p.nb_pad = (n + 7) / 8 * 8;       // tile workgroups keep their XCD order
            nwg = p.nb_pad + ntile;
        }
    }
    if (nwg > 0x7fffffffL) return fail("fd2d_tm_run: %ld workgroups", nwg);
    const dim3 grid((unsigned)nwg);
    if constexpr (PART != TB_WHOLE) {
        // the split launches serve the learned boundary's form only
        if (pml || p.nf) return fail("split TM launch: the form without PML and DFT only");
        if (lossy) hipLaunchKernelGGL((tm_tb_kernel<K, C, true, false, VEC, PART>), grid, dim3(64), 0, s, p);
        else hipLaunchKernelGGL((tm_tb_kernel<K, C, false, false, VEC, PART>), grid, dim3(64), 0, s, p);
        return hip_check(hipGetLastError(), "fd2d_tm_run launch");
    }
    if (p.nf || p.tf_npml) {
        // plane-wave (TF/SF) launches, with or without DFT planes, run the
        // instantiation that carries the planes (nf = 0: none are read or
        // written); the others have no TF/SF code at all, so the waits of
        // their source and probe tiles stay counted
        if (!pml || p.nf > TB_MAX_FREQ)
            return fail("fused plane wave / running DFT: PML form and at most %d frequencies",
                        TB_MAX_FREQ);
        if (lossy)
            hipLaunchKernelGGL((tm_tb_kernel<K, C, true, true, VEC, TB_WHOLE, TB_MAX_FREQ>), grid,
                               dim3(64), 0, s, p);
        else
            hipLaunchKernelGGL((tm_tb_kernel<K, C, false, true, VEC, TB_WHOLE, TB_MAX_FREQ>), grid,
                               dim3(64), 0, s, p);
        return hip_check(hipGetLastError(), "fd2d_tm_run launch");
    }
    if (lossy && pml) hipLaunchKernelGGL((tm_tb_kernel<K, C, true, true, VEC, TB_WHOLE>), grid, dim3(64), 0, s, p);
    else if (lossy) hipLaunchKernelGGL((tm_tb_kernel<K, C, true, false, VEC, TB_WHOLE>), grid, dim3(64), 0, s, p);
    else if (pml) hipLaunchKernelGGL((tm_tb_kernel<K, C, false, true, VEC, TB_WHOLE>), grid, dim3(64), 0, s, p);
    else hipLaunchKernelGGL((tm_tb_kernel<K, C, false, false, VEC, TB_WHOLE>), grid, dim3(64), 0, s, p);
    return hip_check(hipGetLastError(), "fd2d_tm_run launch");
}

template <int K, int C, int PART>
int launch_cols(const TbParams &p, bool lossy, bool pml, hipStream_t s)
{
    // vector loads need every lane's columns wholly inside or outside the grid
    // (the host rounds the halo up to a multiple of the columns per lane)
    if (C == 2 && p.ny % 2 == 0) return launch_kv<K, C, true, PART>(p, lossy, pml, s);
    return launch_kv<K, C, false, PART>(p, lossy, pml, s);
}

template <int K, int C>
int launch_part(const TbParams &p, bool lossy, bool pml, int part, hipStream_t s)
{
    switch (part) {
    case TB_WHOLE: return launch_cols<K, C, TB_WHOLE>(p, lossy, pml, s);
    case TB_INTERIOR: return launch_cols<K, C, TB_INTERIOR>(p, lossy, pml, s);
    case TB_BOUNDARY: return launch_cols<K, C, TB_BOUNDARY>(p, lossy, pml, s);
    default: return fail("TM launch part %d", part);
    }
}

// Whether every naz, nbz, iz (input set) of rows [r0, r1] on the wave's
// columns (j0 .. j0+C-1 per lane, inside the grid) is exactly 1, +0, +0: bit
// patterns, wave-uniform.
template <int C, class PR>
__device__ __forceinline__ bool free_space(const PR &P, int r0, int r1, int j0)
{
    bool ok = true;
#pragma unroll 4
    for (int r = r0; r <= r1; ++r) {
        const unsigned row = (unsigned)r * (unsigned)P.ny + (unsigned)j0;
#pragma unroll
        for (int c = 0; c < C; ++c) {
            ok &= __float_as_uint(ldg(P.naz + row + c)) == 0x3f800000u;
            ok &= __float_as_uint(ldg(P.nbz + row + c)) == 0u;
            ok &= __float_as_uint(ldg(P.iz_i + row + c)) == 0u;
        }
    }
    return __builtin_amdgcn_ballot_w64(!ok) == 0;
}

// ----------------------------------------------------------- batched grids
// A unit's parameters.  The sweep reads the same fields as from TbParams, but
// a batch grid has at most one source and one probe, so the lists are single
// values (One ignores the index) and the plane-wave and DFT terms are absent
// at compile time: nothing in the record is indexed at run time, and it stays
// in registers.  (A TbParams built in the kernel would live in scratch memory,
// since its source and probe lists are indexed by a loop counter.)
struct One {
    int v;
    __device__ int operator[](int) const { return v; }
};

struct BatchParams {
    int nx, ny;
    const float *dz_i, *iz_i, *hx_i, *hy_i, *ihx_i, *ihy_i;
    float *dz_o, *iz_o, *hx_o, *hy_o, *ihx_o, *ihy_o, *ez_o;
    const float *naz, *nbz;
    PmlPtrs pml;
    int t_first;
    int nsrc;
    One src_i, src_j;
    const float *src_wave;
    int wave_len;
    int nprobe;
    One probe_i, probe_j;
    float *probe_out;
    int probe_len;
    unsigned iz_bytes;      // iz of free-space units: 0 (reads +0, no stores) or the grid's bytes
    static constexpr int skip_w = 0;
    static constexpr int tf_npml = 0, inc_t0 = 0, inc_stride = 0;
    static constexpr const float *inc = nullptr;
    static constexpr int nf = 0;
};

// A persistent kernel over a work list of (grid, strip, row block) units of
// many independent grids (fd2d_grid records in device memory); each unit is
// swept exactly as the single-grid kernels sweep it.  Lossy PML form only.
template <int K, int C, bool VEC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(C == 1 && K <= 8 ? 4 : 1)))
tm_tb_batch_kernel(const fd2d_grid *__restrict__ grids, const int *__restrict__ units, int nunits,
                   int parity, int t_first, int write_ez, int halo, BatchVac vac)
{
    constexpr int S = 64 * C;
    __shared__ NaRing<K, C> ring;
    // XCD-aware slots: wave w (on XCD w % 8) takes the units slot, slot +
    // gridDim.x, ...; the slots of one XCD are contiguous, so in each round
    // the neighbouring units of a grid (strips fastest in the list) run on
    // the same XCD and share its L2
    const unsigned nw = gridDim.x, xcd = blockIdx.x % 8u, q = nw / 8u, rem = nw % 8u;
    const unsigned slot = xcd * q + min(xcd, rem) + blockIdx.x / 8u;
    for (int u = (int)slot; u < nunits; u += (int)nw) {
        // unit: grid, strip, rows [i0, i1) (each grid has its own row blocks)
        const int gi = units[4 * u], sx = units[4 * u + 1];
        const int ui0 = units[4 * u + 2], ui1 = units[4 * u + 3];
        const fd2d_grid &G = grids[gi];
        BatchParams P;
        P.nx = G.nx;
        P.ny = G.ny;
        const bool swap = parity != 0;        // odd launches read the partner set
        P.dz_i = swap ? G.dz2 : G.dz;   P.dz_o = swap ? G.dz : G.dz2;
        P.iz_i = swap ? G.iz2 : G.iz;   P.iz_o = swap ? G.iz : G.iz2;
        P.hx_i = swap ? G.hx2 : G.hx;   P.hx_o = swap ? G.hx : G.hx2;
        P.hy_i = swap ? G.hy2 : G.hy;   P.hy_o = swap ? G.hy : G.hy2;
        P.ihx_i = swap ? G.ihx2 : G.ihx; P.ihx_o = swap ? G.ihx : G.ihx2;
        P.ihy_i = swap ? G.ihy2 : G.ihy; P.ihy_o = swap ? G.ihy : G.ihy2;
        P.ez_o = write_ez ? G.ez : nullptr;
        P.naz = G.md.naz;
        P.nbz = G.md.nbz;
        P.pml = PmlPtrs{G.pml.fx1, G.pml.fx2, G.pml.fx3, G.pml.fy1, G.pml.fy2, G.pml.fy3,
                        G.pml.gx2, G.pml.gx3, G.pml.gy2, G.pml.gy3};
        const int *br = G.band_rows, *bc = G.band_cols;
        const int rlo = br[1] > br[0] ? max(br[0], 1) : 0;
        const int rhi = br[1] > br[0] ? min(br[1], G.nx - 1) : 0;
        const int clo = bc[1] > bc[0] ? max(bc[0], 1) : 0;
        const int chi = bc[1] > bc[0] ? min(bc[1], G.ny - 1) : 0;
        P.t_first = t_first;
        P.nsrc = G.src_i >= 0 ? 1 : 0;
        P.src_i.v = G.src_i;
        P.src_j.v = G.src_j;
        P.src_wave = G.src_wave;
        P.wave_len = G.wave_len;
        P.nprobe = G.probe_i >= 0 ? 1 : 0;
        P.probe_i.v = G.probe_i;
        P.probe_j.v = G.probe_j;
        P.probe_out = G.probe_out;
        P.probe_len = G.probe_len;
        const int js = sx * (S - 2 * halo) - halo;
        Geom g;
        g.j0 = js + (int)threadIdx.x * C;
        g.i0 = ui0;
        g.i1 = ui1;
        g.vj0 = js + halo;
        g.vj1 = min(js + S - halo, P.ny);
        const bool brow = !(g.i0 - 2 * K >= rlo && g.i1 - 1 + K < rhi);
        const bool bcol = !(js >= clo && js + S <= chi);
        bool sp = false;
        if (P.nsrc)
            sp |= P.src_i.v >= g.i0 - 2 * K && P.src_i.v <= g.i1 + K && P.src_j.v >= js &&
                  P.src_j.v < js + S;
        if (P.nprobe)
            sp |= P.probe_i.v >= g.i0 && P.probe_i.v < g.i1 && P.probe_j.v >= js &&
                  P.probe_j.v < js + S;
        // free-space units: the first launch of a run call (its deepest)
        // records every unit's entry, the later ones read it.  An interior
        // unit of the recording launch first reads its rows of naz, nbz
        // (those the sweep reads and one more below) and, if they are free
        // space, sweeps without them at once; its traffic is that of a
        // plain sweep.  (Recording inside the sweep, as tm_tb_kernel does,
        // would take registers this kernel does not have at 4 waves per
        // SIMD.)
        if (!brow && !bcol && !sp) {
            bool v = vac.mode == 2 && ldu(vac.work, u) == (vac.tag | 1u);
            // the later launches move no iz in free-space units (+0 in both
            // sets: the recording launch read +0 and writes +0); iz goes
            // through buffer resources of size 0 there, so the recording
            // launch and the later ones share one body (a second free-space
            // body would spill)
            P.iz_bytes = v ? 0u : (unsigned)P.nx * (unsigned)P.ny * 4u;
            if (vac.mode == 1) {
                v = free_space<C>(P, max(g.i0 - K, 0), min(g.i1 + K, P.nx - 1), g.j0);
                if (threadIdx.x == 0) vac.work[u] = vac.tag | (v ? 1u : 0u);
            }
            if (v) sweep<K, C, true, true, VEC, false, false, false, 0, true, C == 1, true, false, 2>(P, g, ring);
            else sweep<K, C, true, true, VEC, false, false, false, 0, true, C == 1>(P, g, ring);
        } else {
            if (vac.mode == 1 && threadIdx.x == 0) vac.work[u] = vac.tag;
            sweep<K, C, true, true, VEC, true, true, true, 0, true>(P, g, ring);
        }
    }
}

template <int K, int C, bool VEC>
int launch_batch_kv(const fd2d_grid *grids, const int *units, int nunits, int waves, int parity,
                    int t_first, int write_ez, int halo, BatchVac vac, hipStream_t s)
{
    hipLaunchKernelGGL((tm_tb_batch_kernel<K, C, VEC>), dim3(waves), dim3(64), 0, s, grids, units,
                       nunits, parity, t_first, write_ez, halo, vac);
    return hip_check(hipGetLastError(), "fd2d_batch_run launch");
}

template <int K, int C, bool VEC>
int batch_resident_kv(int *waves)
{
    int per_cu = 0, dev = 0, ncu = 0;
    int rc = hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                           &per_cu, (const void *)tm_tb_batch_kernel<K, C, VEC>, 64, 0),
                       "occupancy query");
    if (!rc) rc = hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (!rc) rc = hip_check(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev),
                            "CU count");
    if (!rc) *waves = per_cu * ncu;
    return rc;
}

}  // namespace tb

template <int K>
int launch_fd2d_batch_depth(const fd2d_grid *grids, const int *units, int nunits, int waves,
                            int cols, bool vec, int parity, int t_first, int write_ez, int halo,
                            BatchVac vac, hipStream_t s)
{
    if (cols == 2 && vec)
        return tb::launch_batch_kv<K, 2, true>(grids, units, nunits, waves, parity, t_first, write_ez, halo,
                                               vac, s);
    if (cols == 2)
        return tb::launch_batch_kv<K, 2, false>(grids, units, nunits, waves, parity, t_first, write_ez,
                                                halo, vac, s);
    if (cols == 1)
        return tb::launch_batch_kv<K, 1, false>(grids, units, nunits, waves, parity, t_first, write_ez,
                                                halo, vac, s);
    return fail("columns per lane %d not built", cols);
}

template <int K>
int fd2d_batch_resident_depth(int cols, bool vec, int *waves)
{
    if (cols == 2 && vec) return tb::batch_resident_kv<K, 2, true>(waves);
    if (cols == 2) return tb::batch_resident_kv<K, 2, false>(waves);
    if (cols == 1) return tb::batch_resident_kv<K, 1, false>(waves);
    return fail("columns per lane %d not built", cols);
}

template <int K>
int preload_tb_depth()
{
    return preload_kernel(reinterpret_cast<const void *>(&tb::tm_tb_kernel<K, 1, true, true, false, TB_WHOLE>));
}

template <int K>
int launch_fd2d_tb_depth(const TbParams &p, int cols, bool lossy, bool pml, int part, hipStream_t s)
{
    if (cols == 2) return tb::launch_part<K, 2>(p, lossy, pml, part, s);
    if (cols == 1) return tb::launch_part<K, 1>(p, lossy, pml, part, s);
    return fail("columns per lane %d not built", cols);
}

}  // namespace fdtd
