// Instantiation of the temporally blocked TM kernels for depth 1 (fd2d_tb_impl.h).
#include "fd2d_tb_impl.h"

namespace fdtd {
template int launch_fd2d_tb_depth<1>(const TbParams &, int, bool, bool, int, hipStream_t);
}  // namespace fdtd
