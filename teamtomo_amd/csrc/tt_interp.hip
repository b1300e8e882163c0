// Generic 3-D and 2-D sampling and splatting at arbitrary array coordinates
// (the drop-in for torch_image_interpolation's 3-D and 2-D entry points).
//
// sample: packages/primitives/torch-image-interpolation/src/torch_image_interpolation/
//         image_interpolation_3d.py:11-101, image_interpolation_2d.py:11-101
//         (grid_sample, border, align_corners=True, zero outside [0, N-1]);
//         (tri/bi)linear corners floor / floor+1; 2-D bicubic = Keys cubic
//         convolution (A = -0.75) over 4x4 taps clipped to the image.
// splat:  image_interpolation_3d.py:104-275, image_interpolation_2d.py:104-275
//         (index_put_ accumulate); linear corners floor / ceil with float32
//         weights, nearest = round half to even.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"

namespace tt {

namespace {

constexpr int kBlock = 256;

template <typename T> struct Val;
template <> struct Val<float> { typedef float real; typedef f2v cplx; };
template <> struct Val<double> { typedef double real; typedef d2v cplx; };

__device__ __forceinline__ float nearest_index_f(float c, int n) {
  // grid_sample's float round trip: g = c / (0.5n - 0.5) - 1, then
  // u = ((g + 1) / 2) * (n - 1), clipped to [0, n-1], rounded half to even.
  float sc = __fsub_rn(__fmul_rn(0.5f, (float)n), 0.5f);
  float g = __fsub_rn(__fdiv_rn(c, sc), 1.0f);
  float u = __fmul_rn(__fdiv_rn(__fadd_rn(g, 1.0f), 2.0f), (float)(n - 1));
  u = fminf(fmaxf(u, 0.0f), (float)(n - 1));
  return rintf(u);
}

__device__ __forceinline__ double nearest_index_f(double c, int n) {
  double sc = __dsub_rn(__dmul_rn(0.5, (double)n), 0.5);
  double g = __dsub_rn(__ddiv_rn(c, sc), 1.0);
  double u = __dmul_rn(__ddiv_rn(__dadd_rn(g, 1.0), 2.0), (double)(n - 1));
  u = fmin(fmax(u, 0.0), (double)(n - 1));
  return rint(u);
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(kBlock) void sample3d_kernel(const T *__restrict__ img, int C, int d,
                                                          int h, int w,
                                                          const T *__restrict__ coords, int64_t n,
                                                          int mode, T *__restrict__ out) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w, vol = (int64_t)d * plane;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kBlock) {
    T cz = coords[3 * p], cy = coords[3 * p + 1], cx = coords[3 * p + 2];
    bool ok = cz >= 0 && cz <= (T)(d - 1) && cy >= 0 && cy <= (T)(h - 1) && cx >= 0 &&
              cx <= (T)(w - 1);
    T *o = out + p * C * E;
    if (!ok) {
      for (int q = 0; q < C * E; ++q) o[q] = 0;
      continue;
    }
    if (mode == TT_INTERP_NEAREST) {
      int z = (int)nearest_index_f(cz, d), y = (int)nearest_index_f(cy, h),
          x = (int)nearest_index_f(cx, w);
      int64_t e = (int64_t)z * plane + (int64_t)y * w + x;
      for (int ch = 0; ch < C; ++ch)
        for (int q = 0; q < E; ++q) o[ch * E + q] = img[(ch * vol + e) * E + q];
      continue;
    }
    int z0 = (int)floor(cz), y0 = (int)floor(cy), x0 = (int)floor(cx);
    T tz = cz - (T)z0, ty = cy - (T)y0, tx = cx - (T)x0;
    for (int ch = 0; ch < C; ++ch) {
      T acc[2] = {0, 0};
      for (int q = 0; q < 8; ++q) {
        int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
        int z = z0 + dz, y = y0 + dy, x = x0 + dx;
        if (z >= d || y >= h || x >= w) continue;  // weight is 0 there
        T wt = (dz ? tz : (T)1 - tz) * (dy ? ty : (T)1 - ty) * (dx ? tx : (T)1 - tx);
        int64_t e = ch * vol + (int64_t)z * plane + (int64_t)y * w + x;
        for (int k = 0; k < E; ++k) acc[k] += wt * img[e * E + k];
      }
      for (int k = 0; k < E; ++k) o[ch * E + k] = acc[k];
    }
  }
}

template <typename T, typename TW, bool CPLX>
__global__ __launch_bounds__(kBlock) void splat3d_kernel(const T *__restrict__ vals,
                                                         const float *__restrict__ coords,
                                                         int64_t n, int C, int d, int h, int w,
                                                         int mode, T *__restrict__ img,
                                                         TW *__restrict__ wts) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w, vol = (int64_t)d * plane;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kBlock) {
    float cz = coords[3 * p], cy = coords[3 * p + 1], cx = coords[3 * p + 2];
    bool ok = cz >= 0.0f && cz <= (float)(d - 1) && cy >= 0.0f && cy <= (float)(h - 1) &&
              cx >= 0.0f && cx <= (float)(w - 1);
    if (!ok) continue;
    const T *v = vals + p * C * E;
    if (mode == TT_INTERP_NEAREST) {
      int z = (int)rintf(cz), y = (int)rintf(cy), x = (int)rintf(cx);
      int64_t e = (int64_t)z * plane + (int64_t)y * w + x;
      for (int ch = 0; ch < C; ++ch)
        for (int k = 0; k < E; ++k) unsafeAtomicAdd(img + (ch * vol + e) * E + k, v[ch * E + k]);
      unsafeAtomicAdd(wts + e, (TW)1);
      continue;
    }
    int z0 = (int)floorf(cz), y0 = (int)floorf(cy), x0 = (int)floorf(cx);
    int z1 = (int)ceilf(cz), y1 = (int)ceilf(cy), x1 = (int)ceilf(cx);
    float tz = cz - (float)z0, ty = cy - (float)y0, tx = cx - (float)x0;
    for (int q = 0; q < 8; ++q) {
      int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
      int z = dz ? z1 : z0, y = dy ? y1 : y0, x = dx ? x1 : x0;
      float wt = __fmul_rn(__fmul_rn(dz ? tz : 1.0f - tz, dy ? ty : 1.0f - ty), dx ? tx : 1.0f - tx);
      int64_t e = (int64_t)z * plane + (int64_t)y * w + x;
      for (int ch = 0; ch < C; ++ch)
        for (int k = 0; k < E; ++k)
          unsafeAtomicAdd(img + (ch * vol + e) * E + k, v[ch * E + k] * (T)wt);
      unsafeAtomicAdd(wts + e, (TW)wt);
    }
  }
}

// Keys cubic-convolution weights (A = -0.75, grid_sample's bicubic) of the
// taps at -1, 0, 1, 2 for the fraction t.
template <typename T>
__device__ __forceinline__ void cubic_weights(T t, T (&wt)[4]) {
  const T A = (T)-0.75;
  const T x0 = t + (T)1, x1 = t, x2 = (T)1 - t, x3 = (T)2 - t;
  wt[0] = ((A * x0 - (T)5 * A) * x0 + (T)8 * A) * x0 - (T)4 * A;
  wt[1] = ((A + (T)2) * x1 - (A + (T)3)) * x1 * x1 + (T)1;
  wt[2] = ((A + (T)2) * x2 - (A + (T)3)) * x2 * x2 + (T)1;
  wt[3] = ((A * x3 - (T)5 * A) * x3 + (T)8 * A) * x3 - (T)4 * A;
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(kBlock) void sample2d_kernel(const T *__restrict__ img, int C, int h, int w,
                                                          const T *__restrict__ coords, int64_t n,
                                                          int mode, T *__restrict__ out) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kBlock) {
    T cy = coords[2 * p], cx = coords[2 * p + 1];
    bool ok = cy >= 0 && cy <= (T)(h - 1) && cx >= 0 && cx <= (T)(w - 1);
    T *o = out + p * C * E;
    if (!ok) {
      for (int q = 0; q < C * E; ++q) o[q] = 0;
      continue;
    }
    if (mode == TT_INTERP_NEAREST) {
      const int64_t e = (int64_t)nearest_index_f(cy, h) * w + (int64_t)nearest_index_f(cx, w);
      for (int ch = 0; ch < C; ++ch)
        for (int q = 0; q < E; ++q) o[ch * E + q] = img[(ch * plane + e) * E + q];
      continue;
    }
    const int y0 = (int)floor(cy), x0 = (int)floor(cx);
    const T ty = cy - (T)y0, tx = cx - (T)x0;
    if (mode == TT_INTERP_BICUBIC) {
      T wy[4], wx[4];
      cubic_weights(ty, wy);
      cubic_weights(tx, wx);
      for (int ch = 0; ch < C; ++ch) {
        T acc[2] = {0, 0};
        for (int a = 0; a < 4; ++a) {
          const int yy = min(max(y0 - 1 + a, 0), h - 1);
          for (int b = 0; b < 4; ++b) {
            const int xx = min(max(x0 - 1 + b, 0), w - 1);
            const T wt = wy[a] * wx[b];
            const int64_t e = ch * plane + (int64_t)yy * w + xx;
            for (int k = 0; k < E; ++k) acc[k] += wt * img[e * E + k];
          }
        }
        for (int k = 0; k < E; ++k) o[ch * E + k] = acc[k];
      }
      continue;
    }
    for (int ch = 0; ch < C; ++ch) {
      T acc[2] = {0, 0};
      for (int q = 0; q < 4; ++q) {
        const int dy = q >> 1, dx = q & 1;
        const int y = y0 + dy, x = x0 + dx;
        if (y >= h || x >= w) continue;  // weight is 0 there
        const T wt = (dy ? ty : (T)1 - ty) * (dx ? tx : (T)1 - tx);
        const int64_t e = ch * plane + (int64_t)y * w + x;
        for (int k = 0; k < E; ++k) acc[k] += wt * img[e * E + k];
      }
      for (int k = 0; k < E; ++k) o[ch * E + k] = acc[k];
    }
  }
}

template <typename T, typename TW, bool CPLX>
__global__ __launch_bounds__(kBlock) void splat2d_kernel(const T *__restrict__ vals,
                                                         const float *__restrict__ coords,
                                                         int64_t n, int C, int h, int w, int mode,
                                                         T *__restrict__ img, TW *__restrict__ wts) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kBlock) {
    const float cy = coords[2 * p], cx = coords[2 * p + 1];
    if (!(cy >= 0.0f && cy <= (float)(h - 1) && cx >= 0.0f && cx <= (float)(w - 1))) continue;
    const T *v = vals + p * C * E;
    if (mode == TT_INTERP_NEAREST) {
      const int64_t e = (int64_t)rintf(cy) * w + (int64_t)rintf(cx);
      for (int ch = 0; ch < C; ++ch)
        for (int k = 0; k < E; ++k) unsafeAtomicAdd(img + (ch * plane + e) * E + k, v[ch * E + k]);
      unsafeAtomicAdd(wts + e, (TW)1);
      continue;
    }
    const int y0 = (int)floorf(cy), x0 = (int)floorf(cx);
    const int y1 = (int)ceilf(cy), x1 = (int)ceilf(cx);
    const float ty = cy - (float)y0, tx = cx - (float)x0;
    for (int q = 0; q < 4; ++q) {
      const int dy = q >> 1, dx = q & 1;
      const int y = dy ? y1 : y0, x = dx ? x1 : x0;
      const float wt = __fmul_rn(dy ? ty : 1.0f - ty, dx ? tx : 1.0f - tx);
      const int64_t e = (int64_t)y * w + x;
      for (int ch = 0; ch < C; ++ch)
        for (int k = 0; k < E; ++k)
          unsafeAtomicAdd(img + (ch * plane + e) * E + k, v[ch * E + k] * (T)wt);
      unsafeAtomicAdd(wts + e, (TW)wt);
    }
  }
}

int grid_for(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 20));
}

int check_image(int channels, int d, int h, int w) {
  if (channels < 1 || d < 1 || h < 1 || w < 1) return fail_invalid("image dimensions must be >= 1");
  if ((int64_t)channels * d * h * w >= ((int64_t)1 << 40)) return fail_invalid("image too large");
  return TT_OK;
}

template <typename T>
int sample_impl(const T *image, int channels, int d, int h, int w, int is_complex, const T *coords,
                int64_t n, int mode, T *out, void *stream) {
  int st = check_image(channels, d, h, w);
  if (st) return st;
  if (mode != TT_INTERP_NEAREST && mode != TT_INTERP_TRILINEAR)
    return fail_invalid("unknown interpolation mode");
  if (n <= 0) return TT_OK;
  if (!image || !coords || !out) return fail_invalid("NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (is_complex)
    hipLaunchKernelGGL((sample3d_kernel<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, s, image,
                       channels, d, h, w, coords, n, mode, out);
  else
    hipLaunchKernelGGL((sample3d_kernel<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, s, image,
                       channels, d, h, w, coords, n, mode, out);
  return check_launch("sample3d_kernel");
}

template <typename T>
int splat_impl(const T *values, const float *coords, int64_t n, int channels, int is_complex, int d,
               int h, int w, int mode, T *image, void *weights, int weights_f64, void *stream) {
  int st = check_image(channels, d, h, w);
  if (st) return st;
  if (mode != TT_INTERP_NEAREST && mode != TT_INTERP_TRILINEAR)
    return fail_invalid("unknown interpolation mode");
  if (n <= 0) return TT_OK;
  if (!values || !coords || !image || !weights) return fail_invalid("NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  dim3 grid(grid_for(n)), block(kBlock);
  if (weights_f64) {
    if (is_complex)
      hipLaunchKernelGGL((splat3d_kernel<T, double, true>), grid, block, 0, s, values, coords, n,
                         channels, d, h, w, mode, image, (double *)weights);
    else
      hipLaunchKernelGGL((splat3d_kernel<T, double, false>), grid, block, 0, s, values, coords, n,
                         channels, d, h, w, mode, image, (double *)weights);
  } else {
    if (is_complex)
      hipLaunchKernelGGL((splat3d_kernel<T, float, true>), grid, block, 0, s, values, coords, n,
                         channels, d, h, w, mode, image, (float *)weights);
    else
      hipLaunchKernelGGL((splat3d_kernel<T, float, false>), grid, block, 0, s, values, coords, n,
                         channels, d, h, w, mode, image, (float *)weights);
  }
  return check_launch("splat3d_kernel");
}

template <typename T>
int sample2d_impl(const T *image, int channels, int h, int w, int is_complex, const T *coords, int64_t n,
                  int mode, T *out, void *stream) {
  int st = check_image(channels, 1, h, w);
  if (st) return st;
  if (mode != TT_INTERP_NEAREST && mode != TT_INTERP_BILINEAR && mode != TT_INTERP_BICUBIC)
    return fail_invalid("unknown interpolation mode");
  if (n <= 0) return TT_OK;
  if (!image || !coords || !out) return fail_invalid("NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (is_complex)
    hipLaunchKernelGGL((sample2d_kernel<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, s, image, channels,
                       h, w, coords, n, mode, out);
  else
    hipLaunchKernelGGL((sample2d_kernel<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, s, image, channels,
                       h, w, coords, n, mode, out);
  return check_launch("sample2d_kernel");
}

template <typename T>
int splat2d_impl(const T *values, const float *coords, int64_t n, int channels, int is_complex, int h,
                 int w, int mode, T *image, void *weights, int weights_f64, void *stream) {
  int st = check_image(channels, 1, h, w);
  if (st) return st;
  if (mode != TT_INTERP_NEAREST && mode != TT_INTERP_BILINEAR)
    return fail_invalid("unknown interpolation mode");
  if (n <= 0) return TT_OK;
  if (!values || !coords || !image || !weights) return fail_invalid("NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  dim3 grid(grid_for(n)), block(kBlock);
  if (weights_f64) {
    if (is_complex)
      hipLaunchKernelGGL((splat2d_kernel<T, double, true>), grid, block, 0, s, values, coords, n, channels,
                         h, w, mode, image, (double *)weights);
    else
      hipLaunchKernelGGL((splat2d_kernel<T, double, false>), grid, block, 0, s, values, coords, n, channels,
                         h, w, mode, image, (double *)weights);
  } else {
    if (is_complex)
      hipLaunchKernelGGL((splat2d_kernel<T, float, true>), grid, block, 0, s, values, coords, n, channels,
                         h, w, mode, image, (float *)weights);
    else
      hipLaunchKernelGGL((splat2d_kernel<T, float, false>), grid, block, 0, s, values, coords, n, channels,
                         h, w, mode, image, (float *)weights);
  }
  return check_launch("splat2d_kernel");
}

}  // namespace

}  // namespace tt

using namespace tt;

extern "C" {

int tt_sample3d_f32(const float *image, int channels, int d, int h, int w, int is_complex,
                    const float *coords, int64_t n, int mode, float *out, void *stream) {
  return sample_impl<float>(image, channels, d, h, w, is_complex, coords, n, mode, out, stream);
}

int tt_sample3d_f64(const double *image, int channels, int d, int h, int w, int is_complex,
                    const double *coords, int64_t n, int mode, double *out, void *stream) {
  return sample_impl<double>(image, channels, d, h, w, is_complex, coords, n, mode, out, stream);
}

int tt_splat3d_f32(const float *values, const float *coords, int64_t n, int channels,
                   int is_complex, int d, int h, int w, int mode, float *image, void *weights,
                   int weights_f64, void *stream) {
  return splat_impl<float>(values, coords, n, channels, is_complex, d, h, w, mode, image, weights,
                           weights_f64, stream);
}

int tt_splat3d_f64(const double *values, const float *coords, int64_t n, int channels,
                   int is_complex, int d, int h, int w, int mode, double *image, void *weights,
                   int weights_f64, void *stream) {
  return splat_impl<double>(values, coords, n, channels, is_complex, d, h, w, mode, image, weights,
                            weights_f64, stream);
}

int tt_sample2d_f32(const float *image, int channels, int h, int w, int is_complex, const float *coords,
                    int64_t n, int mode, float *out, void *stream) {
  return sample2d_impl<float>(image, channels, h, w, is_complex, coords, n, mode, out, stream);
}

int tt_sample2d_f64(const double *image, int channels, int h, int w, int is_complex, const double *coords,
                    int64_t n, int mode, double *out, void *stream) {
  return sample2d_impl<double>(image, channels, h, w, is_complex, coords, n, mode, out, stream);
}

int tt_splat2d_f32(const float *values, const float *coords, int64_t n, int channels, int is_complex,
                   int h, int w, int mode, float *image, void *weights, int weights_f64, void *stream) {
  return splat2d_impl<float>(values, coords, n, channels, is_complex, h, w, mode, image, weights,
                             weights_f64, stream);
}

int tt_splat2d_f64(const double *values, const float *coords, int64_t n, int channels, int is_complex,
                   int h, int w, int mode, double *image, void *weights, int weights_f64, void *stream) {
  return splat2d_impl<double>(values, coords, n, channels, is_complex, h, w, mode, image, weights,
                              weights_f64, stream);
}

}  // extern "C"
