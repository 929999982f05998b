// Launches of the warp-specialised stage kernel on meshes with Dirichlet
// faces.  Compiled apart from solver.hip: the kernel instantiations without
// Dirichlet faces (the common case) then compile exactly as they do alone,
// without the inlining choices a second instantiation in the same
// translation unit would change.
#include "kernels/dg_ws.hpp"

namespace dglm {

template <int N, int MODE>
static void launch_one(int grid, hipStream_t stream, const StageArgs &a, const double *afrag, const double *arem,
                       const int *order, int nbatch, int groups) {
  if constexpr (ws_available<N>())
    hipLaunchKernelGGL((stage_ws_kernel<N, MODE, true>), dim3(grid), dim3(WsShape<N>::NT), 0, stream, a, afrag,
                       arem, order, nbatch, groups);
}

bool launch_ws_dirichlet(int N, int mode, int grid, hipStream_t stream, const StageArgs &a, const double *afrag,
                         const double *arem, const int *order, int nbatch, int groups) {
  switch (N * 2 + mode) {
    case 3 * 2 + 0:
      launch_one<3, 0>(grid, stream, a, afrag, arem, order, nbatch, groups);
      return true;
    case 3 * 2 + 1:
      launch_one<3, 1>(grid, stream, a, afrag, arem, order, nbatch, groups);
      return true;
    case 4 * 2 + 0:
      launch_one<4, 0>(grid, stream, a, afrag, arem, order, nbatch, groups);
      return true;
    case 4 * 2 + 1:
      launch_one<4, 1>(grid, stream, a, afrag, arem, order, nbatch, groups);
      return true;
    default:
      return false;
  }
}

}  // namespace dglm
