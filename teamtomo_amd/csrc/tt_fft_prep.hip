// The elementwise steps around the FFTs of project_3d_to_2d and
// backproject_2d_to_3d, each fused into one pass over the volume.
//
// Reference: packages/primitives/torch-fourier-slice/src/torch_fourier_slice/
//   project.py:44-70: pad by p, multiply by sinc^2(|f|) on the centred grid
//   f_i = (i - N/2)/N (fftshift of fftfreq, torch_grid_utils fftfreq_grid), then
//   fftshift all axes before the rfft;
//   backproject.py:75-91: divide the inserted rfft by clamp(weights, 1),
//   ifftshift the two full axes, and after the inverse rfft ifftshift all axes,
//   divide by the same sinc^2 grid and crop the padding.
// The FFTs stay on torch.  Each kernel reads its input once and writes its
// output once (torch runs pad, the sinc^2 grid, the product and the shifts as
// separate passes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"

namespace tt {

namespace {

constexpr int kPrepBlock = 256;

int prep_grid(int64_t n) {
  const int64_t b = (n + kPrepBlock - 1) / kPrepBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20));
}

// sinc(|f|)^2, f_i = (i - N/2) / N on each axis of an N^3 grid (FP64).
__device__ __forceinline__ double sinc2_at(int z, int y, int x, int N) {
  const int h = N / 2;
  const double fz = (double)(z - h) / N, fy = (double)(y - h) / N, fx = (double)(x - h) / N;
  const double r = sqrt(fz * fz + fy * fy + fx * fx);
  if (r == 0.0) return 1.0;
  const double a = 3.14159265358979323846 * r;
  const double s = sin(a) / a;
  return s * s;
}

// out (N^3, N = d + 2 pad) = fftshift(pad(vol) * sinc^2): output index o holds
// input index k = (o + N - N/2) mod N on each axis.
template <typename T>
__global__ __launch_bounds__(kPrepBlock) void pad_sinc2_shift_kernel(const T *__restrict__ vol, int d, int pad,
                                                                    T *__restrict__ out) {
  const int N = d + 2 * pad, sh = N - N / 2;
  const int64_t total = (int64_t)N * N * N;
  for (int64_t q = (int64_t)blockIdx.x * kPrepBlock + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * kPrepBlock) {
    const int oz = (int)(q / ((int64_t)N * N)), rem = (int)(q - (int64_t)oz * N * N);
    const int oy = rem / N, ox = rem - oy * N;
    const int kz = (oz + sh) % N, ky = (oy + sh) % N, kx = (ox + sh) % N;
    const int z = kz - pad, y = ky - pad, x = kx - pad;
    T v = 0;
    if (z >= 0 && z < d && y >= 0 && y < d && x >= 0 && x < d)
      v = (T)((double)vol[((int64_t)z * d + y) * d + x] * sinc2_at(kz, ky, kx, N));
    out[q] = v;
  }
}

// out (d^3) = crop(ifftshift(x) op sinc^2), x (N^3): output (z, y, x) reads
// x at (i + N/2) mod N of i = z + pad (and so on); op = divide (the
// backprojection's gridding correction) or multiply (the adjoint of
// pad_sinc2_shift_kernel).
template <typename T, bool DIVIDE>
__global__ __launch_bounds__(kPrepBlock) void unshift_sinc2_crop_kernel(const T *__restrict__ x, int N, int pad,
                                                                       T *__restrict__ out) {
  const int d = N - 2 * pad, sh = N / 2;
  const int64_t total = (int64_t)d * d * d;
  for (int64_t q = (int64_t)blockIdx.x * kPrepBlock + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * kPrepBlock) {
    const int z = (int)(q / ((int64_t)d * d)), rem = (int)(q - (int64_t)z * d * d);
    const int y = rem / d, xx = rem - y * d;
    const int iz = z + pad, iy = y + pad, ix = xx + pad;
    const double s = sinc2_at(iz, iy, ix, N);
    const double v = (double)x[((int64_t)((iz + sh) % N) * N + (iy + sh) % N) * N + (ix + sh) % N];
    out[q] = (T)(DIVIDE ? v / s : v * s);
  }
}

// out = ifftshift over the two full axes of vol / clamp(w, 1), (D, H, Wr).
template <typename C, typename W>
__global__ __launch_bounds__(kPrepBlock) void wdiv_unshift_kernel(const C *__restrict__ vol, const W *__restrict__ w,
                                                                 int D, int H, int Wr, C *__restrict__ out) {
  const int64_t total = (int64_t)D * H * Wr;
  for (int64_t q = (int64_t)blockIdx.x * kPrepBlock + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * kPrepBlock) {
    const int oz = (int)(q / ((int64_t)H * Wr)), rem = (int)(q - (int64_t)oz * H * Wr);
    const int oy = rem / Wr, ox = rem - oy * Wr;
    const int64_t e = ((int64_t)((oz + D / 2) % D) * H + (oy + H / 2) % H) * Wr + ox;
    const W den = w[e] > (W)1 ? w[e] : (W)1;
    C v = vol[e];
    v.x /= den;
    v.y /= den;
    out[q] = v;
  }
}

template <typename T>
int pad_impl(const T *vol, int d, int pad, T *out, void *stream) {
  if (d < 1 || pad < 0) return fail_invalid("volume side must be >= 1 and padding >= 0");
  if (!vol || !out) return fail_invalid("NULL pointer");
  const int N = d + 2 * pad;
  hipLaunchKernelGGL(pad_sinc2_shift_kernel<T>, dim3(prep_grid((int64_t)N * N * N)), dim3(kPrepBlock), 0,
                     (hipStream_t)stream, vol, d, pad, out);
  return check_launch("pad_sinc2_shift_kernel");
}

template <typename T>
int unpad_impl(const T *x, int N, int pad, int divide, T *out, void *stream) {
  if (N < 1 || pad < 0 || N - 2 * pad < 1) return fail_invalid("padded side must exceed twice the padding");
  if (!x || !out) return fail_invalid("NULL pointer");
  const int d = N - 2 * pad;
  const dim3 grid(prep_grid((int64_t)d * d * d)), block(kPrepBlock);
  if (divide)
    hipLaunchKernelGGL((unshift_sinc2_crop_kernel<T, true>), grid, block, 0, (hipStream_t)stream, x, N, pad, out);
  else
    hipLaunchKernelGGL((unshift_sinc2_crop_kernel<T, false>), grid, block, 0, (hipStream_t)stream, x, N, pad, out);
  return check_launch("unshift_sinc2_crop_kernel");
}

template <typename C, typename W>
int wdiv_impl(const C *vol, const W *w, int D, int H, int Wr, C *out, void *stream) {
  if (D < 1 || H < 1 || Wr < 1) return fail_invalid("volume dimensions must be >= 1");
  if (!vol || !w || !out) return fail_invalid("NULL pointer");
  hipLaunchKernelGGL((wdiv_unshift_kernel<C, W>), dim3(prep_grid((int64_t)D * H * Wr)), dim3(kPrepBlock), 0,
                     (hipStream_t)stream, vol, w, D, H, Wr, out);
  return check_launch("wdiv_unshift_kernel");
}

}  // namespace

}  // namespace tt

using namespace tt;

extern "C" {

int tt_pad_sinc2_fftshift_f32(const float *vol, int d, int pad, float *out, void *stream) {
  return pad_impl<float>(vol, d, pad, out, stream);
}
int tt_pad_sinc2_fftshift_f64(const double *vol, int d, int pad, double *out, void *stream) {
  return pad_impl<double>(vol, d, pad, out, stream);
}
int tt_ifftshift_sinc2_crop_f32(const float *x, int N, int pad, int divide, float *out, void *stream) {
  return unpad_impl<float>(x, N, pad, divide, out, stream);
}
int tt_ifftshift_sinc2_crop_f64(const double *x, int N, int pad, int divide, double *out, void *stream) {
  return unpad_impl<double>(x, N, pad, divide, out, stream);
}
int tt_wdiv_ifftshift_c64(const void *vol, const float *w, int D, int H, int Wr, void *out, void *stream) {
  return wdiv_impl<f2v, float>((const f2v *)vol, w, D, H, Wr, (f2v *)out, stream);
}
int tt_wdiv_ifftshift_c128(const void *vol, const double *w, int D, int H, int Wr, void *out, void *stream) {
  return wdiv_impl<d2v, double>((const d2v *)vol, w, D, H, Wr, (d2v *)out, stream);
}

}  // extern "C"
