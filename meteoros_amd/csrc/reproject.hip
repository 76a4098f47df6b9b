// reproject.hip -- 4x4 temporal reprojection update (kernel c):
// reprojection.comp main (:193-244) per pixel, the arithmetic in
// reproject_pixel.h.  One lane per pixel; a workgroup of 256 lanes covers a
// 64x4 pixel strip so each wave reads one contiguous 64-pixel row of the
// previous image per tap where the motion is small.
#include "reproject_pixel.h"

namespace mtc {
namespace reproj {

// The plain frames' reprojection (every pixel; the march then overwrites
// its 1/16) and the tap-index map.  The optimised frames carry theirs in
// the march launch (march_fast.hip), leaving the marched pixels out.
template <bool INDICES>
__global__ __launch_bounds__(256) void reproject_kernel(FrameConsts fc, OutputViews ov) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= fc.width || y >= fc.height) return;
  reproject_pixel<INDICES>(fc, ov, x, y);
}

}  // namespace reproj
}  // namespace mtc

// mode 0: the update; 1: the tap-index map
hipError_t launch_reproject(const FrameConsts& fc, const OutputViews& ov, int mode, hipStream_t stream,
                            const KernelEvents& ke) {
  if (mode != 0 && mode != 1) return hipErrorInvalidValue;
  dim3 grid((fc.width + 63) / 64, (fc.height + 3) / 4);
  dim3 block(256);
  if (mode == 1) launch_timed(mtc::reproj::reproject_kernel<true>, grid, block, stream, ke, fc, ov);
  else launch_timed(mtc::reproj::reproject_kernel<false>, grid, block, stream, ke, fc, ov);
  return hipGetLastError();
}
