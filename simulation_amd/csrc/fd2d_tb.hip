// fd2d_tb.hip -- depth dispatch of the temporally blocked TM kernels
// (kernel body: fd2d_tb_impl.h; one translation unit per depth,
// fd2d_tb_k<depth>.hip, so the variants compile in parallel).
#include "fdtd_internal.h"

namespace fdtd {

int tb_max_depth() { return TB_MAX_DEPTH; }

int preload_tb()
{
    int rc = 0;
    rc = rc ? rc : preload_tb_depth<1>();
    rc = rc ? rc : preload_tb_depth<2>();
    rc = rc ? rc : preload_tb_depth<3>();
    rc = rc ? rc : preload_tb_depth<4>();
    rc = rc ? rc : preload_tb_depth<5>();
    rc = rc ? rc : preload_tb_depth<6>();
    rc = rc ? rc : preload_tb_depth<7>();
    rc = rc ? rc : preload_tb_depth<8>();
    rc = rc ? rc : preload_tb_depth<9>();
    rc = rc ? rc : preload_tb_depth<10>();
    rc = rc ? rc : preload_tb_depth<11>();
    rc = rc ? rc : preload_tb_depth<12>();
    return rc;
}

int launch_fd2d_mw(const TbParams &p, int depth, bool lossy, bool pml, int rows, int frows,
                   hipStream_t s)
{
    switch (depth) {
    case 1: return launch_fd2d_mw_depth<1>(p, lossy, pml, rows, frows, s);
    case 2: return launch_fd2d_mw_depth<2>(p, lossy, pml, rows, frows, s);
    case 3: return launch_fd2d_mw_depth<3>(p, lossy, pml, rows, frows, s);
    case 4: return launch_fd2d_mw_depth<4>(p, lossy, pml, rows, frows, s);
    case 5: return launch_fd2d_mw_depth<5>(p, lossy, pml, rows, frows, s);
    case 6: return launch_fd2d_mw_depth<6>(p, lossy, pml, rows, frows, s);
    case 7: return launch_fd2d_mw_depth<7>(p, lossy, pml, rows, frows, s);
    case 8: return launch_fd2d_mw_depth<8>(p, lossy, pml, rows, frows, s);
    default: return fail("temporal depth %d not built", depth);
    }
}

int preload_mw()
{
    int rc = 0;
    rc = rc ? rc : preload_mw_depth<1>();
    rc = rc ? rc : preload_mw_depth<2>();
    rc = rc ? rc : preload_mw_depth<3>();
    rc = rc ? rc : preload_mw_depth<4>();
    rc = rc ? rc : preload_mw_depth<5>();
    rc = rc ? rc : preload_mw_depth<6>();
    rc = rc ? rc : preload_mw_depth<7>();
    rc = rc ? rc : preload_mw_depth<8>();
    return rc;
}

bool tb_supported(int depth, int cols)
{
    return depth >= 1 && depth <= tb_max_depth() && (cols == 1 || cols == 2);
}

int launch_fd2d_tb(const TbParams &p, int depth, int cols, bool lossy, bool pml, int part,
                   hipStream_t s)
{
    switch (depth) {
    case 1: return launch_fd2d_tb_depth<1>(p, cols, lossy, pml, part, s);
    case 2: return launch_fd2d_tb_depth<2>(p, cols, lossy, pml, part, s);
    case 3: return launch_fd2d_tb_depth<3>(p, cols, lossy, pml, part, s);
    case 4: return launch_fd2d_tb_depth<4>(p, cols, lossy, pml, part, s);
    case 5: return launch_fd2d_tb_depth<5>(p, cols, lossy, pml, part, s);
    case 6: return launch_fd2d_tb_depth<6>(p, cols, lossy, pml, part, s);
    case 7: return launch_fd2d_tb_depth<7>(p, cols, lossy, pml, part, s);
    case 8: return launch_fd2d_tb_depth<8>(p, cols, lossy, pml, part, s);
    case 9: return launch_fd2d_tb_depth<9>(p, cols, lossy, pml, part, s);
    case 10: return launch_fd2d_tb_depth<10>(p, cols, lossy, pml, part, s);
    case 11: return launch_fd2d_tb_depth<11>(p, cols, lossy, pml, part, s);
    case 12: return launch_fd2d_tb_depth<12>(p, cols, lossy, pml, part, s);
    default: return fail("temporal depth %d not built", depth);
    }
}

int launch_fd2d_batch(const fd2d_grid *grids, const int *units, int nunits, int waves, int depth,
                      int cols, bool vec, int parity, int t_first, int write_ez, int halo,
                      BatchVac vac, hipStream_t s)
{
#define FDTD_BATCH_CASE(k)                                                                   \
    case k:                                                                                  \
        return launch_fd2d_batch_depth<k>(grids, units, nunits, waves, cols, vec, parity,    \
                                          t_first, write_ez, halo, vac, s);
    switch (depth) {
        FDTD_BATCH_CASE(1) FDTD_BATCH_CASE(2) FDTD_BATCH_CASE(3) FDTD_BATCH_CASE(4)
        FDTD_BATCH_CASE(5) FDTD_BATCH_CASE(6) FDTD_BATCH_CASE(7) FDTD_BATCH_CASE(8)
        FDTD_BATCH_CASE(9) FDTD_BATCH_CASE(10) FDTD_BATCH_CASE(11) FDTD_BATCH_CASE(12)
    default: return fail("temporal depth %d not built", depth);
    }
#undef FDTD_BATCH_CASE
}

int fd2d_batch_resident(int depth, int cols, bool vec, int *waves)
{
    switch (depth) {
    case 1: return fd2d_batch_resident_depth<1>(cols, vec, waves);
    case 2: return fd2d_batch_resident_depth<2>(cols, vec, waves);
    case 3: return fd2d_batch_resident_depth<3>(cols, vec, waves);
    case 4: return fd2d_batch_resident_depth<4>(cols, vec, waves);
    case 5: return fd2d_batch_resident_depth<5>(cols, vec, waves);
    case 6: return fd2d_batch_resident_depth<6>(cols, vec, waves);
    case 7: return fd2d_batch_resident_depth<7>(cols, vec, waves);
    case 8: return fd2d_batch_resident_depth<8>(cols, vec, waves);
    case 9: return fd2d_batch_resident_depth<9>(cols, vec, waves);
    case 10: return fd2d_batch_resident_depth<10>(cols, vec, waves);
    case 11: return fd2d_batch_resident_depth<11>(cols, vec, waves);
    case 12: return fd2d_batch_resident_depth<12>(cols, vec, waves);
    default: return fail("temporal depth %d not built", depth);
    }
}

}  // namespace fdtd
