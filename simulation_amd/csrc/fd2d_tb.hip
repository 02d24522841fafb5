// fd2d_tb.hip -- depth dispatch of the temporally blocked TM kernels
// (kernel body: fd2d_tb_impl.h; one translation unit per depth,
// fd2d_tb_k<depth>.hip, so the variants compile in parallel).
#include "fdtd_internal.h"

namespace fdtd {

int tb_max_depth() { return 8; }

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
    default: return fail("temporal depth %d not built", depth);
    }
}

}  // namespace fdtd
