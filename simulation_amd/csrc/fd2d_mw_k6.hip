// Instantiation of the multi-wave temporally blocked TM kernel for depth 6 (fd2d_mw_impl.h).
#include "fd2d_mw_impl.h"

namespace fdtd {
template int launch_fd2d_mw_depth<6>(const TbParams &, bool, bool, int, int, hipStream_t);
template int preload_mw_depth<6>();
}  // namespace fdtd
