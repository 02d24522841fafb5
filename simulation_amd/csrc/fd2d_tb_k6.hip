// Instantiation of the temporally blocked TM kernels for depth 6 (fd2d_tb_impl.h).
#include "fd2d_tb_impl.h"

namespace fdtd {
template int launch_fd2d_tb_depth<6>(const TbParams &, int, bool, bool, int, hipStream_t);
}  // namespace fdtd
