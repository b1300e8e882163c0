// Fourier shifts (torch-fourier-shift drop-ins): a batch of DFTs times the
// phase ramp exp(-2 pi i f . s) of each image's shift s, in one pass, and the
// backward of that product (the DFT gradient and the shift gradient) in one
// more.
//
// Reference: packages/primitives/torch-fourier-shift/src/torch_fourier_shift/
//   phase_shift_grids.py:9-127: angle = -2 pi sum_a f_a s_a with f from
//   fftfreq on the full axes and rfftfreq (rfft) or fftfreq on the last
//   (fftfreq_grids.py:9-116), then fftshift over every axis, or every axis but
//   the last for an rfft (dft_utils.py:12-54);
//   fourier_shift_dft.py:10-115: the DFT times that grid.
// The ramp is evaluated per element: the phase in cycles, a = sum_a f_a s_a,
// is formed in FP64 and reduced to [-1/2, 1/2] before the sine and cosine
// (sincospi), so its error does not grow with |s| as a float32 angle's does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"

namespace tt {

namespace {

constexpr int kShiftBlock = 256;
constexpr int kMaxGridY = 65535;
constexpr int kGradBlocksPerImage = 64;

// Layout of one batch entry: (d, h, ws) stored elements (d = h = 1 for lower
// dimensionality), image sizes (D, Hh, Wf) of the axes, ws = Wf/2 + 1 for an
// rfft.
struct ShiftGeom {
  int ndim;        // 1, 2 or 3
  int d, h, ws;    // stored extents (leading ones 1 for lower ndim)
  int D, Hh, Wf;   // image extents along z, y, x
  int rfft, fftshifted;
  int per;         // d * h * ws
  double inv_D, inv_H, inv_W;  // 1/D, 1/Hh, 1/Wf
  double inv_hw, inv_ws;       // 1/(h ws), 1/ws
};

// floor(v / d) for 0 <= v < 2^31 without an integer division: the FP64
// product with 1/d is within one of the quotient, and the remainder fixes it.
__device__ __forceinline__ int div_fix(int v, int d, double inv_d) {
  int q = (int)((double)v * inv_d);
  const int r = v - q * d;
  if (r < 0)
    --q;
  else if (r >= d)
    ++q;
  return q;
}

// fftfreq value of stored index k on a full axis of size n (inv_n = 1/n).
__device__ __forceinline__ double full_freq(int k, int n, double inv_n, int shifted) {
  const int m = shifted ? k - n / 2 : (k < (n + 1) / 2 ? k : k - n);
  return (double)m * inv_n;
}

// Frequencies (fz, fy, fx) of stored element v of one batch entry.
__device__ __forceinline__ void shift_freqs(const ShiftGeom &g, int v, double &fz, double &fy, double &fx) {
  const int hw = g.h * g.ws;
  const int kz = g.ndim == 3 ? div_fix(v, hw, g.inv_hw) : 0, rem = v - kz * hw;
  const int ky = div_fix(rem, g.ws, g.inv_ws), kx = rem - ky * g.ws;
  fz = g.ndim == 3 ? full_freq(kz, g.D, g.inv_D, g.fftshifted) : 0.0;
  fy = g.ndim >= 2 ? full_freq(ky, g.Hh, g.inv_H, g.fftshifted) : 0.0;
  fx = g.rfft ? (double)kx * g.inv_W : full_freq(kx, g.Wf, g.inv_W, g.fftshifted);
}

// Phase in cycles, reduced to [-1/2, 1/2].
template <typename S>
__device__ __forceinline__ double shift_cycles(const ShiftGeom &g, const S *s, double fz, double fy, double fx) {
  double a = fx * (double)s[g.ndim - 1];
  if (g.ndim >= 2) a = fma(fy, (double)s[g.ndim - 2], a);
  if (g.ndim == 3) a = fma(fz, (double)s[0], a);
  return a - rint(a);
}

// exp(-2 pi i a) as (cos, sin).
__device__ __forceinline__ f2v ramp(float, double a) {
  float sn, cs;
  sincospif(-2.0f * (float)a, &sn, &cs);
  return f2v{cs, sn};
}
__device__ __forceinline__ d2v ramp(double, double a) {
  double sn, cs;
  sincospi(-2.0 * a, &sn, &cs);
  return d2v{cs, sn};
}

template <typename T>
__device__ __forceinline__ typename Cx<T>::type cmul(typename Cx<T>::type a, typename Cx<T>::type b) {
  return typename Cx<T>::type{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// out[t, v] = dft[t * dft_stride + v] * exp(-2 pi i f(v) . s[t]); blockIdx.y
// walks the batch (a uniform shift per block), blockIdx.x the elements.
// dft and out may be the same buffer (each element is read, then written, by
// one thread).
template <typename T>
__global__ __launch_bounds__(kShiftBlock) void fourier_shift_kernel(const typename Cx<T>::type *dft,
                                                                    int64_t dft_stride, const T *__restrict__ shifts,
                                                                    int64_t n, ShiftGeom g,
                                                                    typename Cx<T>::type *out) {
  for (int64_t t = blockIdx.y; t < n; t += gridDim.y) {
    const T *s = shifts + t * g.ndim;
    const typename Cx<T>::type *src = dft + t * dft_stride;
    typename Cx<T>::type *dst = out + t * (int64_t)g.per;
    for (int v = blockIdx.x * kShiftBlock + threadIdx.x; v < g.per; v += gridDim.x * kShiftBlock) {
      double fz, fy, fx;
      shift_freqs(g, v, fz, fy, fx);
      dst[v] = cmul<T>(src[v], ramp(T(), shift_cycles(g, s, fz, fy, fx)));
    }
  }
}

// Backward of out = dft * ramp(s): grad_dft = g * conj(ramp) (written when
// grad_dft is set) and, when partial is set, per block the three sums
// sum_v f_a(v) Im(conj(g) out) over its elements, which the finalize kernel
// turns into grad_s_a = 2 pi sum.  Reads g[t, v] and dft[t * dft_stride + v].
template <typename T>
__global__ __launch_bounds__(kShiftBlock) void fourier_shift_bwd_kernel(
    const typename Cx<T>::type *__restrict__ grad, const typename Cx<T>::type *__restrict__ dft, int64_t dft_stride,
    const T *__restrict__ shifts, int64_t n, ShiftGeom g, typename Cx<T>::type *__restrict__ grad_dft,
    double *__restrict__ partial) {
  __shared__ double red[3][kShiftBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t t = blockIdx.y; t < n; t += gridDim.y) {
    const T *s = shifts + t * g.ndim;
    const typename Cx<T>::type *gs = grad + t * (int64_t)g.per;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int v = blockIdx.x * kShiftBlock + threadIdx.x; v < g.per; v += gridDim.x * kShiftBlock) {
      double fz, fy, fx;
      shift_freqs(g, v, fz, fy, fx);
      const typename Cx<T>::type e = ramp(T(), shift_cycles(g, s, fz, fy, fx));
      const typename Cx<T>::type gv = gs[v];
      if (grad_dft) {
        const typename Cx<T>::type ce{e.x, -e.y};
        grad_dft[t * (int64_t)g.per + v] = cmul<T>(gv, ce);
      }
      if (partial) {
        const typename Cx<T>::type y = cmul<T>(dft[t * dft_stride + v], e);
        const double c = (double)gv.x * (double)y.y - (double)gv.y * (double)y.x;  // Im(conj(g) y)
        acc[0] = fma(fz, c, acc[0]);
        acc[1] = fma(fy, c, acc[1]);
        acc[2] = fma(fx, c, acc[2]);
      }
    }
    if (partial) {  // block-uniform branch
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        double x = acc[a];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[a][wave] = x;
      }
      __syncthreads();
      if (threadIdx.x < 3) {
        double x = 0.0;
        for (int w = 0; w < kShiftBlock / 64; ++w) x += red[threadIdx.x][w];
        partial[(t * gridDim.x + blockIdx.x) * 3 + threadIdx.x] = x;
      }
      __syncthreads();  // red is rewritten for the next batch entry
    }
  }
}

// grad_s[t, a] = 2 pi sum_b partial[t, b, axis(a)], in block order (the result
// does not depend on the schedule).
template <typename T>
__global__ __launch_bounds__(kShiftBlock) void fourier_shift_grad_finalize(const double *__restrict__ partial,
                                                                          int64_t n, int nblk, int ndim,
                                                                          T *__restrict__ grad_shifts) {
  const int64_t q = (int64_t)blockIdx.x * kShiftBlock + threadIdx.x;
  if (q >= n * ndim) return;
  const int64_t t = q / ndim;
  const int a = (int)(q - t * ndim);
  const int axis = 3 - ndim + a;  // shifts are (z, y, x) trailing: partial slot 0 z, 1 y, 2 x
  double x = 0.0;
  for (int b = 0; b < nblk; ++b) x += partial[(t * nblk + b) * 3 + axis];
  grad_shifts[q] = (T)(6.283185307179586476925286766559 * x);
}

int make_shift_geom(int ndim, int d, int h, int w, int rfft, int fftshifted, ShiftGeom &g) {
  if (ndim < 1 || ndim > 3) return fail_invalid("ndim must be 1, 2 or 3");
  if (w < 1 || (ndim >= 2 && h < 1) || (ndim == 3 && d < 1)) return fail_invalid("image dimensions must be >= 1");
  g.ndim = ndim;
  g.D = ndim == 3 ? d : 1;
  g.Hh = ndim >= 2 ? h : 1;
  g.Wf = w;
  g.d = g.D;
  g.h = g.Hh;
  g.ws = rfft ? w / 2 + 1 : w;
  g.rfft = rfft ? 1 : 0;
  g.fftshifted = fftshifted ? 1 : 0;
  const int64_t per = (int64_t)g.d * g.h * g.ws;
  if (per >= (int64_t)1 << 31) return fail_invalid("one DFT must hold fewer than 2^31 elements");
  g.per = (int)per;
  g.inv_D = 1.0 / g.D;
  g.inv_H = 1.0 / g.Hh;
  g.inv_W = 1.0 / g.Wf;
  g.inv_hw = 1.0 / ((double)g.h * g.ws);
  g.inv_ws = 1.0 / g.ws;
  return TT_OK;
}

dim3 shift_grid(const ShiftGeom &g, int64_t n, int max_x) {
  const int bx = (int)std::min<int64_t>(max_x, std::max<int64_t>(1, (g.per + kShiftBlock - 1) / kShiftBlock));
  return dim3(bx, (unsigned)std::min<int64_t>(n, kMaxGridY));
}

template <typename T>
int shift_fwd(const void *dft, int64_t dft_stride, const T *shifts, int64_t n, int ndim, int d, int h, int w,
              int rfft, int fftshifted, void *out, void *stream) {
  ShiftGeom g;
  if (int st = make_shift_geom(ndim, d, h, w, rfft, fftshifted, g)) return st;
  if (dft_stride != 0 && dft_stride != g.per) return fail_invalid("dft_stride must be 0 or one DFT's element count");
  if (n <= 0) return TT_OK;
  if (!dft || !shifts || !out) return fail_invalid("NULL pointer");
  typedef typename Cx<T>::type C;
  hipLaunchKernelGGL(fourier_shift_kernel<T>, shift_grid(g, n, 1 << 16), dim3(kShiftBlock), 0, (hipStream_t)stream,
                     (const C *)dft, dft_stride, shifts, n, g, (C *)out);
  return check_launch("fourier_shift_kernel");
}

template <typename T>
int shift_bwd(const void *grad, const void *dft, int64_t dft_stride, const T *shifts, int64_t n, int ndim, int d,
              int h, int w, int rfft, int fftshifted, void *grad_dft, T *grad_shifts, void *stream, const char *name) {
  ShiftGeom g;
  if (int st = make_shift_geom(ndim, d, h, w, rfft, fftshifted, g)) return st;
  if (dft_stride != 0 && dft_stride != g.per) return fail_invalid("dft_stride must be 0 or one DFT's element count");
  if (n <= 0 || (!grad_dft && !grad_shifts)) return TT_OK;
  if (!grad || !shifts || (grad_shifts && !dft)) return fail_invalid("NULL pointer");
  typedef typename Cx<T>::type C;
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid = shift_grid(g, n, kGradBlocksPerImage);
  double *partial = nullptr;
  if (grad_shifts) {
    if (hipError_t e = hipMallocAsync((void **)&partial, (size_t)n * grid.x * 3 * sizeof(double), s)) {
      set_error(std::string(name) + ": hipMallocAsync: " + hipGetErrorString(e));
      return TT_ERR_HIP;
    }
  }
  hipLaunchKernelGGL(fourier_shift_bwd_kernel<T>, grid, dim3(kShiftBlock), 0, s, (const C *)grad, (const C *)dft,
                     dft_stride, shifts, n, g, (C *)grad_dft, partial);
  int st = check_launch("fourier_shift_bwd_kernel");
  if (grad_shifts) {
    if (!st) {
      const int64_t q = n * ndim;
      hipLaunchKernelGGL(fourier_shift_grad_finalize<T>, dim3((unsigned)((q + kShiftBlock - 1) / kShiftBlock)),
                         dim3(kShiftBlock), 0, s, (const double *)partial, n, (int)grid.x, ndim, grad_shifts);
      st = check_launch("fourier_shift_grad_finalize");
    }
    if (hipError_t e = hipFreeAsync(partial, s)) {
      if (!st) {
        set_error(std::string(name) + ": hipFreeAsync: " + hipGetErrorString(e));
        st = TT_ERR_HIP;
      }
    }
  }
  return st;
}

}  // namespace

}  // namespace tt

using namespace tt;

extern "C" {

int tt_fourier_shift_c64(const void *dft, int64_t dft_stride, const float *shifts, int64_t n, int ndim, int d, int h,
                         int w, int rfft, int fftshifted, void *out, void *stream) {
  return shift_fwd<float>(dft, dft_stride, shifts, n, ndim, d, h, w, rfft, fftshifted, out, stream);
}

int tt_fourier_shift_c128(const void *dft, int64_t dft_stride, const double *shifts, int64_t n, int ndim, int d,
                          int h, int w, int rfft, int fftshifted, void *out, void *stream) {
  return shift_fwd<double>(dft, dft_stride, shifts, n, ndim, d, h, w, rfft, fftshifted, out, stream);
}

int tt_fourier_shift_bwd_c64(const void *grad_out, const void *dft, int64_t dft_stride, const float *shifts,
                             int64_t n, int ndim, int d, int h, int w, int rfft, int fftshifted, void *grad_dft,
                             float *grad_shifts, void *stream) {
  return shift_bwd<float>(grad_out, dft, dft_stride, shifts, n, ndim, d, h, w, rfft, fftshifted, grad_dft,
                          grad_shifts, stream, "tt_fourier_shift_bwd_c64");
}

int tt_fourier_shift_bwd_c128(const void *grad_out, const void *dft, int64_t dft_stride, const double *shifts,
                              int64_t n, int ndim, int d, int h, int w, int rfft, int fftshifted, void *grad_dft,
                              double *grad_shifts, void *stream) {
  return shift_bwd<double>(grad_out, dft, dft_stride, shifts, n, ndim, d, h, w, rfft, fftshifted, grad_dft,
                           grad_shifts, stream, "tt_fourier_shift_bwd_c128");
}

}  // extern "C"
