// reproject.hip -- 4x4 temporal reprojection update (kernel c):
// reprojection.comp main (:193-244) per pixel, the arithmetic in
// reproject_pixel.h.  One lane per pixel; a workgroup of 256 lanes covers a
// 64x4 pixel strip so each wave reads one contiguous 64-pixel row of the
// previous image per tap where the motion is small.
#include "reproject_pixel.h"

namespace mtc {
namespace reproj {

// SKIP_SELECTED: leave out the pixels this frame's march overwrites (the
// 4x4 phase (frame / 4, frame % 4)), so the two kernels write disjoint
// pixels and may run concurrently.
template <bool INDICES, bool SKIP_SELECTED>
__global__ __launch_bounds__(256) void reproject_kernel(FrameConsts fc, OutputViews ov) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= fc.width || y >= fc.height) return;
  if (SKIP_SELECTED && marched_this_frame(fc, x, y)) return;
  reproject_pixel<INDICES>(fc, ov, x, y);
}

}  // namespace reproj
}  // namespace mtc

// mode 0: the update; 1: the tap-index map; 2: the update without the
// pixels the march of this frame writes
hipError_t launch_reproject(const FrameConsts& fc, const OutputViews& ov, int mode, hipStream_t stream,
                            const KernelEvents& ke) {
  dim3 grid((fc.width + 63) / 64, (fc.height + 3) / 4);
  dim3 block(256);
  if (mode == 1) launch_timed(mtc::reproj::reproject_kernel<true, false>, grid, block, stream, ke, fc, ov);
  else if (mode == 2) launch_timed(mtc::reproj::reproject_kernel<false, true>, grid, block, stream, ke, fc, ov);
  else launch_timed(mtc::reproj::reproject_kernel<false, false>, grid, block, stream, ke, fc, ov);
  return hipGetLastError();
}
