// Instantiation of the temporally blocked TM kernels (single grid and batched) for depth 3 (fd2d_tb_impl.h).
#include "fd2d_tb_impl.h"

namespace fdtd {
template int launch_fd2d_tb_depth<3>(const TbParams &, int, bool, bool, int, hipStream_t);
template int launch_fd2d_batch_depth<3>(const fd2d_grid *, const int *, int, int, int, bool, int, int,
                                        int, int, BatchVac, hipStream_t);
template int fd2d_batch_resident_depth<3>(int, bool, int *);
template int preload_tb_depth<3>();
}  // namespace fdtd
