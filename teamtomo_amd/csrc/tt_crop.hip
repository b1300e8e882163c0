// Subpixel crops (torch-subpixel-crop drop-ins): an integer-position crop with
// zero fill and an optional real-space mask.  The Fourier phase shift that
// moves each patch by its subpixel remainder is tt_fourier_shift_c64.
//
// Reference: packages/primitives/torch-subpixel-crop/src/torch_subpixel_crop/
//   subpixel_crop_3d.py:12-108, subpixel_crop_2d.py:12-163: patches of
//   sidelength p around round(position) (grid_sample nearest at integer
//   coordinates, zeros outside; the patch centre is index p//2), times the
//   mask, then torch-fourier-shift's phase ramp exp(-2 pi i f . s) on the
//   rfft (packages/primitives/torch-fourier-shift/src/torch_fourier_shift/
//   phase_shift_grids.py:60-127, f from fftfreq / rfftfreq, not fftshifted).
// The FFTs stay on torch; the crop is one memory pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"

namespace tt {

namespace {

constexpr int kCropBlock = 256;

int crop_grid(int64_t n) {
  const int64_t b = (n + kCropBlock - 1) / kCropBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20));
}

// out[m, b, k...] = image[b, pos[m, b] + k - p/2] (0 outside) * mask[k...];
// NDIM = 2 or 3, images (B, [D,] H, W), positions (n, B, NDIM) int32.
template <int NDIM>
__global__ __launch_bounds__(kCropBlock) void crop_kernel(const float *__restrict__ img, int B, int D, int H, int W,
                                                         const int *__restrict__ pos, int64_t n, int p,
                                                         const float *__restrict__ mask, float *__restrict__ out) {
  const int64_t per = NDIM == 3 ? (int64_t)p * p * p : (int64_t)p * p;
  const int64_t total = n * B * per;
  const int c = p / 2;
  for (int64_t q = (int64_t)blockIdx.x * kCropBlock + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * kCropBlock) {
    const int64_t mb = q / per;  // (m, b) flattened
    const int k = (int)(q - mb * per);
    const int b = (int)(mb % B);
    const int *ps = pos + mb * NDIM;
    int kz = 0, ky, kx;
    if (NDIM == 3) {
      kz = k / (p * p);
      ky = (k / p) % p;
    } else {
      ky = k / p;
    }
    kx = k % p;
    const int z = NDIM == 3 ? ps[0] + kz - c : 0;
    const int y = ps[NDIM - 2] + ky - c, x = ps[NDIM - 1] + kx - c;
    float v = 0.0f;
    if ((NDIM == 2 || (z >= 0 && z < D)) && y >= 0 && y < H && x >= 0 && x < W)
      v = img[(((int64_t)b * D + z) * H + y) * W + x];
    if (mask) v *= mask[k];
    out[q] = v;
  }
}

}  // namespace

}  // namespace tt

using namespace tt;

extern "C" {

int tt_crop3d_f32(const float *image, int D, int H, int W, const int *positions, int64_t n, int sidelength,
                  const float *mask, float *patches, void *stream) {
  if (D < 1 || H < 1 || W < 1 || sidelength < 1) return fail_invalid("image and patch sizes must be >= 1");
  if (n <= 0) return TT_OK;
  if (!image || !positions || !patches) return fail_invalid("NULL pointer");
  const int64_t total = n * (int64_t)sidelength * sidelength * sidelength;
  hipLaunchKernelGGL(crop_kernel<3>, dim3(crop_grid(total)), dim3(kCropBlock), 0, (hipStream_t)stream, image, 1, D,
                     H, W, positions, n, sidelength, mask, patches);
  return check_launch("crop_kernel<3>");
}

int tt_crop2d_f32(const float *images, int B, int H, int W, const int *positions, int64_t n, int sidelength,
                  const float *mask, float *patches, void *stream) {
  if (B < 1 || H < 1 || W < 1 || sidelength < 1) return fail_invalid("image and patch sizes must be >= 1");
  if (n <= 0) return TT_OK;
  if (!images || !positions || !patches) return fail_invalid("NULL pointer");
  const int64_t total = n * B * (int64_t)sidelength * sidelength;
  hipLaunchKernelGGL(crop_kernel<2>, dim3(crop_grid(total)), dim3(kCropBlock), 0, (hipStream_t)stream, images, B, 1,
                     H, W, positions, n, sidelength, mask, patches);
  return check_launch("crop_kernel<2>");
}

}  // extern "C"
