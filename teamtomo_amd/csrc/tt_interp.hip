// Generic 3-D and 2-D sampling and splatting at arbitrary array coordinates
// (the drop-in for torch_image_interpolation's 3-D and 2-D entry points).
//
// sample: packages/primitives/torch-image-interpolation/src/torch_image_interpolation/
//         image_interpolation_3d.py:11-101, image_interpolation_2d.py:11-101
//         (grid_sample, border, align_corners=True, zero outside [0, N-1]);
//         (tri/bi)linear corners floor / floor+1; 2-D bicubic = Keys cubic
//         convolution (A = -0.75) over 4x4 taps clipped to the image.
// affine: torch-transform-image .../transforms_3d.py:10-40, transforms_2d.py:11-41
//         (the sampler at M (z, y, x, 1)^T of every output voxel, fused).
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

// One 3-D sample (all channels) at array coordinate (cz, cy, cx) into o.
template <typename T, bool CPLX>
__device__ __forceinline__ void sample3d_point(const T *__restrict__ img, int C, int d, int h, int w, T cz,
                                               T cy, T cx, int mode, T *__restrict__ o) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w, vol = (int64_t)d * plane;
  const bool ok = cz >= 0 && cz <= (T)(d - 1) && cy >= 0 && cy <= (T)(h - 1) && cx >= 0 && cx <= (T)(w - 1);
  if (!ok) {
    for (int q = 0; q < C * E; ++q) o[q] = 0;
    return;
  }
  if (mode == TT_INTERP_NEAREST) {
    int z = (int)nearest_index_f(cz, d), y = (int)nearest_index_f(cy, h), x = (int)nearest_index_f(cx, w);
    int64_t e = (int64_t)z * plane + (int64_t)y * w + x;
    for (int ch = 0; ch < C; ++ch)
      for (int q = 0; q < E; ++q) o[ch * E + q] = img[(ch * vol + e) * E + q];
    return;
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

template <typename T, bool CPLX>
__global__ __launch_bounds__(kBlock) void sample3d_kernel(const T *__restrict__ img, int C, int d,
                                                          int h, int w,
                                                          const T *__restrict__ coords, int64_t n,
                                                          int mode, T *__restrict__ out) {
  const int E = CPLX ? 2 : 1;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kBlock)
    sample3d_point<T, CPLX>(img, C, d, h, w, coords[3 * p], coords[3 * p + 1], coords[3 * p + 2], mode,
                            out + p * C * E);
}

// grid_sample's source coordinate of array coordinate c on an axis of n
// samples, in the sampler's precision: the normalisation of
// grid_sample_utils.py:24, g = c / (0.5 n - 0.5) - 1, undone as
// u = ((g + 1) / 2) (n - 1) (align_corners=True), unclipped.  The round trip
// can move an integer coordinate by an ulp, which decides the cell (and so
// the one-sided derivative) there.
__device__ __forceinline__ float gs_source(float c, int n) {
  if (n <= 1) return 0.0f;
  float sc = __fsub_rn(__fmul_rn(0.5f, (float)n), 0.5f);
  float g = __fsub_rn(__fdiv_rn(c, sc), 1.0f);
  return __fmul_rn(__fdiv_rn(__fadd_rn(g, 1.0f), 2.0f), (float)(n - 1));
}

__device__ __forceinline__ double gs_source(double c, int n) {
  if (n <= 1) return 0.0;
  double sc = __dsub_rn(__dmul_rn(0.5, (double)n), 0.5);
  double g = __dsub_rn(__ddiv_rn(c, sc), 1.0);
  return __dmul_rn(__ddiv_rn(__dadd_rn(g, 1.0), 2.0), (double)(n - 1));
}

// Backward of sample3d_point for one sample: the image gradient (trilinear
// weights of the floor / floor+1 corners, or the nearest voxel, added with
// global atomics into a zeroed gimg) and the coordinate gradient,
// Re(conj(g) ds/dc), from the analytic derivative of the weights.  As in
// grid_sample's backward with border padding and align_corners
// (image_interpolation_3d.py:64-75), the weights are those of the source
// coordinate u (gs_source) clipped to [0, N-1], the derivative along an axis
// is 0 where the clip holds it (u <= 0 or u >= N-1), nearest sampling has no
// coordinate gradient, and samples outside the image (zeroed in the forward,
// :84-89) get no gradient at all.
template <typename T, bool CPLX>
__device__ __forceinline__ void sample3d_point_bwd(const T *__restrict__ img, int C, int d, int h, int w, T cz,
                                                   T cy, T cx, int mode, const T *__restrict__ g,
                                                   T *__restrict__ gimg, T (&gc)[3]) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w, vol = (int64_t)d * plane;
  gc[0] = gc[1] = gc[2] = 0;
  const bool ok = cz >= 0 && cz <= (T)(d - 1) && cy >= 0 && cy <= (T)(h - 1) && cx >= 0 && cx <= (T)(w - 1);
  if (!ok) return;
  if (mode == TT_INTERP_NEAREST) {
    if (!gimg) return;
    int z = (int)nearest_index_f(cz, d), y = (int)nearest_index_f(cy, h), x = (int)nearest_index_f(cx, w);
    int64_t e = (int64_t)z * plane + (int64_t)y * w + x;
    for (int ch = 0; ch < C; ++ch)
      for (int k = 0; k < E; ++k) atomicAdd(&gimg[(ch * vol + e) * E + k], g[ch * E + k]);
    return;
  }
  const T uz = gs_source(cz, d), uy = gs_source(cy, h), ux = gs_source(cx, w);
  const bool clz = uz <= 0 || uz >= (T)(d - 1), cly = uy <= 0 || uy >= (T)(h - 1),
             clx = ux <= 0 || ux >= (T)(w - 1);
  const T vz = fmin(fmax(uz, (T)0), (T)(d - 1)), vy = fmin(fmax(uy, (T)0), (T)(h - 1)),
          vx = fmin(fmax(ux, (T)0), (T)(w - 1));
  int z0 = (int)floor(vz), y0 = (int)floor(vy), x0 = (int)floor(vx);
  T tz = vz - (T)z0, ty = vy - (T)y0, tx = vx - (T)x0;
  for (int q = 0; q < 8; ++q) {
    int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
    int z = z0 + dz, y = y0 + dy, x = x0 + dx;
    if (z >= d || y >= h || x >= w) continue;  // weight and derivatives are 0 there
    const T wz = dz ? tz : (T)1 - tz, wy = dy ? ty : (T)1 - ty, wx = dx ? tx : (T)1 - tx;
    const T sz = dz ? (T)1 : (T)-1, sy = dy ? (T)1 : (T)-1, sx = dx ? (T)1 : (T)-1;
    const T wt = wz * wy * wx;
    const T dwz = sz * wy * wx, dwy = wz * sy * wx, dwx = wz * wy * sx;
    for (int ch = 0; ch < C; ++ch) {
      int64_t e = ch * vol + (int64_t)z * plane + (int64_t)y * w + x;
      for (int k = 0; k < E; ++k) {
        const T gk = g[ch * E + k], v = img[e * E + k];
        if (gimg) atomicAdd(&gimg[e * E + k], wt * gk);
        gc[0] += gk * v * dwz;
        gc[1] += gk * v * dwy;
        gc[2] += gk * v * dwx;
      }
    }
  }
  if (clz) gc[0] = 0;
  if (cly) gc[1] = 0;
  if (clx) gc[2] = 0;
}

template <typename T, bool CPLX>
__global__ __launch_bounds__(kBlock) void sample3d_bwd_kernel(const T *__restrict__ img, int C, int d, int h,
                                                              int w, const T *__restrict__ coords, int64_t n,
                                                              int mode, const T *__restrict__ gout,
                                                              T *__restrict__ gimg, T *__restrict__ gcoords) {
  const int E = CPLX ? 2 : 1;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n; p += (int64_t)gridDim.x * kBlock) {
    T gc[3];
    sample3d_point_bwd<T, CPLX>(img, C, d, h, w, coords[3 * p], coords[3 * p + 1], coords[3 * p + 2], mode,
                                gout + p * C * E, gimg, gc);
    if (gcoords) {
      gcoords[3 * p] = gc[0];
      gcoords[3 * p + 1] = gc[1];
      gcoords[3 * p + 2] = gc[2];
    }
  }
}

// Fused affine warp (affine_transform_image_3d): output voxel (z, y, x) of
// leading batch index t samples the image at M (z, y, x, 1)^T, M the 3x4 zyx
// block (row-major, float32, as the reference's float32 matmul of the
// coordinate grid) of matrix ((t md + z') mh + y') mw + x', where z' = z if
// md > 1 else 0 (and so on): the matrices broadcast against (d, h, w) as in
// the reference's matmul.  No coordinate grid is materialised.
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void affine3d_kernel(const float *__restrict__ img, int C, int D, int H,
                                                          int W, const float *__restrict__ mats, int64_t n_lead,
                                                          int md, int mh, int mw, int d, int h, int w, int mode,
                                                          float *__restrict__ out) {
  const int E = CPLX ? 2 : 1;
  const int64_t per = (int64_t)d * h * w, total = n_lead * per;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * kBlock) {
    const int64_t t = p / per;
    const int64_t v = p - t * per;
    const int z = (int)(v / ((int64_t)h * w)), rem = (int)(v - (int64_t)z * h * w);
    const int y = rem / w, x = rem - y * w;
    const int64_t mi = ((t * md + (md > 1 ? z : 0)) * mh + (mh > 1 ? y : 0)) * mw + (mw > 1 ? x : 0);
    const float *M = mats + mi * 12;
    const float fz = (float)z, fy = (float)y, fx = (float)x;
    float c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
      c[i] = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(M[4 * i], fz), __fmul_rn(M[4 * i + 1], fy)),
                                 __fmul_rn(M[4 * i + 2], fx)), M[4 * i + 3]);
    sample3d_point<float, CPLX>(img, C, D, H, W, c[0], c[1], c[2], mode, out + p * C * E);
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

// One 2-D sample (all channels) at array coordinate (cy, cx) into o.
template <typename T, bool CPLX>
__device__ __forceinline__ void sample2d_point(const T *__restrict__ img, int C, int h, int w, T cy, T cx,
                                               int mode, T *__restrict__ o) {
  const int E = CPLX ? 2 : 1;
  const int64_t plane = (int64_t)h * w;
  const bool ok = cy >= 0 && cy <= (T)(h - 1) && cx >= 0 && cx <= (T)(w - 1);
  if (!ok) {
    for (int q = 0; q < C * E; ++q) o[q] = 0;
    return;
  }
  if (mode == TT_INTERP_NEAREST) {
    const int64_t e = (int64_t)nearest_index_f(cy, h) * w + (int64_t)nearest_index_f(cx, w);
    for (int ch = 0; ch < C; ++ch)
      for (int q = 0; q < E; ++q) o[ch * E + q] = img[(ch * plane + e) * E + q];
    return;
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
    return;
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

template <typename T, bool CPLX>
__global__ __launch_bounds__(kBlock) void sample2d_kernel(const T *__restrict__ img, int C, int h, int w,
                                                          const T *__restrict__ coords, int64_t n,
                                                          int mode, T *__restrict__ out) {
  const int E = CPLX ? 2 : 1;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < n;
       p += (int64_t)gridDim.x * kBlock)
    sample2d_point<T, CPLX>(img, C, h, w, coords[2 * p], coords[2 * p + 1], mode, out + p * C * E);
}

// Fused affine warp (affine_transform_image_2d): output pixel (y, x) of
// leading batch index t samples the image at M (y, x, 1)^T, M the 2x3 yx
// block of matrix (t mh + y') mw + x' (broadcast as in the 3-D warp).
template <bool CPLX>
__global__ __launch_bounds__(kBlock) void affine2d_kernel(const float *__restrict__ img, int C, int H, int W,
                                                          const float *__restrict__ mats, int64_t n_lead, int mh,
                                                          int mw, int h, int w, int mode, float *__restrict__ out) {
  const int E = CPLX ? 2 : 1;
  const int64_t per = (int64_t)h * w, total = n_lead * per;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * kBlock) {
    const int64_t t = p / per;
    const int v = (int)(p - t * per);
    const int y = v / w, x = v - y * w;
    const int64_t mi = (t * mh + (mh > 1 ? y : 0)) * mw + (mw > 1 ? x : 0);
    const float *M = mats + mi * 6;
    const float fy = (float)y, fx = (float)x;
    const float cy = __fadd_rn(__fadd_rn(__fmul_rn(M[0], fy), __fmul_rn(M[1], fx)), M[2]);
    const float cx = __fadd_rn(__fadd_rn(__fmul_rn(M[3], fy), __fmul_rn(M[4], fx)), M[5]);
    sample2d_point<float, CPLX>(img, C, H, W, cy, cx, mode, out + p * C * E);
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
int sample_bwd_impl(const T *image, int channels, int d, int h, int w, int is_complex, const T *coords,
                    int64_t n, int mode, const T *grad_out, T *grad_image, T *grad_coords, void *stream) {
  int st = check_image(channels, d, h, w);
  if (st) return st;
  if (mode != TT_INTERP_NEAREST && mode != TT_INTERP_TRILINEAR)
    return fail_invalid("unknown interpolation mode");
  if (n <= 0 || (!grad_image && !grad_coords)) return TT_OK;
  if (!image || !coords || !grad_out) return fail_invalid("NULL pointer");
  hipStream_t s = (hipStream_t)stream;
  if (is_complex)
    hipLaunchKernelGGL((sample3d_bwd_kernel<T, true>), dim3(grid_for(n)), dim3(kBlock), 0, s, image,
                       channels, d, h, w, coords, n, mode, grad_out, grad_image, grad_coords);
  else
    hipLaunchKernelGGL((sample3d_bwd_kernel<T, false>), dim3(grid_for(n)), dim3(kBlock), 0, s, image,
                       channels, d, h, w, coords, n, mode, grad_out, grad_image, grad_coords);
  return check_launch("sample3d_bwd_kernel");
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

int affine_impl(int dims, const float *image, int channels, const int *in_shape, int is_complex, const float *mats,
                int64_t n_lead, const int *mshape, const int *out_shape, int mode, float *out, void *stream) {
  for (int i = 0; i < dims; ++i) {
    if (in_shape[i] < 1 || out_shape[i] < 1) return fail_invalid("image and output dimensions must be >= 1");
    if (mshape[i] != 1 && mshape[i] != out_shape[i])
      return fail_invalid("matrix batch dimensions must be 1 or the output size (broadcasting)");
  }
  if (channels < 1) return fail_invalid("channels must be >= 1");
  if (mode != TT_INTERP_NEAREST && mode != TT_INTERP_TRILINEAR && !(dims == 2 && mode == TT_INTERP_BICUBIC))
    return fail_invalid("unknown interpolation mode");
  if (n_lead <= 0) return TT_OK;
  if (!image || !mats || !out) return fail_invalid("NULL pointer");
  int64_t per = 1;
  for (int i = 0; i < dims; ++i) per *= out_shape[i];
  if (per >= ((int64_t)1 << 40) || n_lead > ((int64_t)1 << 40) / per) return fail_invalid("output too large");
  const int64_t total = n_lead * per;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(grid_for(total)), block(kBlock);
  if (dims == 3) {
    if (is_complex)
      hipLaunchKernelGGL(affine3d_kernel<true>, grid, block, 0, s, image, channels, in_shape[0], in_shape[1],
                         in_shape[2], mats, n_lead, mshape[0], mshape[1], mshape[2], out_shape[0], out_shape[1],
                         out_shape[2], mode, out);
    else
      hipLaunchKernelGGL(affine3d_kernel<false>, grid, block, 0, s, image, channels, in_shape[0], in_shape[1],
                         in_shape[2], mats, n_lead, mshape[0], mshape[1], mshape[2], out_shape[0], out_shape[1],
                         out_shape[2], mode, out);
    return check_launch("affine3d_kernel");
  }
  if (is_complex)
    hipLaunchKernelGGL(affine2d_kernel<true>, grid, block, 0, s, image, channels, in_shape[0], in_shape[1], mats,
                       n_lead, mshape[0], mshape[1], out_shape[0], out_shape[1], mode, out);
  else
    hipLaunchKernelGGL(affine2d_kernel<false>, grid, block, 0, s, image, channels, in_shape[0], in_shape[1], mats,
                       n_lead, mshape[0], mshape[1], out_shape[0], out_shape[1], mode, out);
  return check_launch("affine2d_kernel");
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

int tt_sample3d_bwd_f32(const float *image, int channels, int d, int h, int w, int is_complex,
                        const float *coords, int64_t n, int mode, const float *grad_out, float *grad_image,
                        float *grad_coords, void *stream) {
  return sample_bwd_impl<float>(image, channels, d, h, w, is_complex, coords, n, mode, grad_out, grad_image,
                                grad_coords, stream);
}

int tt_sample3d_bwd_f64(const double *image, int channels, int d, int h, int w, int is_complex,
                        const double *coords, int64_t n, int mode, const double *grad_out, double *grad_image,
                        double *grad_coords, void *stream) {
  return sample_bwd_impl<double>(image, channels, d, h, w, is_complex, coords, n, mode, grad_out, grad_image,
                                 grad_coords, stream);
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

int tt_affine3d_f32(const float *image, int channels, int D, int H, int W, int is_complex, const float *mats,
                    int64_t n_lead, int md, int mh, int mw, int d, int h, int w, int mode, float *out,
                    void *stream) {
  const int in_shape[3] = {D, H, W}, mshape[3] = {md, mh, mw}, out_shape[3] = {d, h, w};
  return affine_impl(3, image, channels, in_shape, is_complex, mats, n_lead, mshape, out_shape, mode, out, stream);
}

int tt_affine2d_f32(const float *image, int channels, int H, int W, int is_complex, const float *mats,
                    int64_t n_lead, int mh, int mw, int h, int w, int mode, float *out, void *stream) {
  const int in_shape[2] = {H, W}, mshape[2] = {mh, mw}, out_shape[2] = {h, w};
  return affine_impl(2, image, channels, in_shape, is_complex, mats, n_lead, mshape, out_shape, mode, out, stream);
}

}  // extern "C"
