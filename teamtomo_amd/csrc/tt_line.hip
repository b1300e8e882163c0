// Central lines of 2-D rffts: the 2-D analogue of the slice gather, and its
// backward (image gradient by the adjoint splat, rotation gradient from the
// bilinear-weight derivatives).
//
// Reference: packages/primitives/torch-fourier-slice/src/torch_fourier_slice/
//   slice_extraction/_extract_central_slices_rfft_2d.py:9-92: the line
//   (0, j/W), j < W//2+1 (_grids/_central_line_fftfreq_grid.py:6-31), rotated
//   by R_yx, negated and read conjugated where x' < 0, mapped to array
//   coordinates (f_y + 0.5)(h-1)/(1-1/h) and 2 f_x (Wr-1) (_dft_utils.py:46-67)
//   and sampled bilinearly with zero outside the image array
//   (torch-image-interpolation .../image_interpolation_2d.py:11-101).
//
// Lines are short (W//2+1 pixels) and read a small image that stays in L2, so
// one workgroup takes one rotation; no LDS staging is needed.  Coordinates are
// evaluated in FP64 and snapped within 1e-9 of an integer, as in the 3-D path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"

namespace tt {

namespace {

constexpr int kLineBlock = 256;

struct LineGeom {
  int H, WR;          // image array: rows, rfft columns
  int wr;             // line length W//2+1 of the image shape
  double ay, by, ax;  // array units per cycle/pixel (y full, x half), y of f = 0
  double inv_W;       // 1/W of the image shape
  int use_mask;       // fftfreq_max given
  float fmax, inv_W32;
};

LineGeom make_line_geom(int h, int wr_arr, int img_h, int img_w, double fftfreq_max) {
  LineGeom g;
  g.H = h;
  g.WR = wr_arr;
  g.wr = img_w / 2 + 1;
  g.ay = (double)(img_h - 1) / (1.0 - 1.0 / (double)img_h);
  g.by = 0.5 * g.ay;
  g.ax = 2.0 * (double)(g.wr - 1);
  g.inv_W = 1.0 / (double)img_w;
  g.use_mask = fftfreq_max >= 0.0;
  g.fmax = (float)fftfreq_max;
  g.inv_W32 = (float)(1.0 / (double)img_w);
  return g;
}

struct LineCoord {
  double y, x;
  bool conj;
};

// Rotated, Friedel-folded, snapped array coordinate of line pixel j under the
// yx rotation R (row-major 2x2): y' = R[0][1] f, x' = R[1][1] f, f = j/W.
__device__ __forceinline__ LineCoord line_coord(const LineGeom &g, double Vy, double Vx, int j) {
  const double jd = (double)j;
  const double cx1 = __dmul_rn(jd, Vx);
  LineCoord c;
  c.conj = cx1 < 0.0;
  const double s = c.conj ? -1.0 : 1.0;
  c.y = snap(__fma_rn(jd, s * Vy, g.by));
  c.x = snap(c.conj ? -cx1 : cx1);
  return c;
}

// The line's fftfreq_max cut: |f| = f_x = float32(j) * float32(1/W) <= fmax.
__device__ __forceinline__ bool line_ok(const LineGeom &g, const LineCoord &c, int j) {
  if (g.use_mask && !(__fmul_rn((float)j, g.inv_W32) <= g.fmax)) return false;
  return c.y >= 0.0 && c.y <= (double)(g.H - 1) && c.x >= 0.0 && c.x <= (double)(g.WR - 1);
}

// Lower corner (clamped so the upper corner is addressable when the axis has
// two or more samples) and fraction.
__device__ __forceinline__ void line_cell(double c, int n, int &i0, float &t, bool &up) {
  i0 = (int)c;
  if (i0 > n - 2) i0 = n >= 2 ? n - 2 : 0;
  t = (float)(c - (double)i0);
  up = i0 + 1 < n;
}

__device__ __forceinline__ void line_rot(const LineGeom &g, const float *__restrict__ R, double &Vy,
                                         double &Vx) {
  Vy = __dmul_rn(__dmul_rn(g.ay, (double)R[1]), g.inv_W);
  Vx = __dmul_rn(__dmul_rn(g.ax, (double)R[3]), g.inv_W);
}

__global__ __launch_bounds__(kLineBlock) void line_fwd_kernel(const f2v *__restrict__ img, LineGeom g,
                                                             const float *__restrict__ rot,
                                                             f2v *__restrict__ out) {
  const int64_t r = blockIdx.x;
  double Vy, Vx;
  line_rot(g, rot + r * 4, Vy, Vx);
  for (int j = threadIdx.x; j < g.wr; j += kLineBlock) {
    const LineCoord c = line_coord(g, Vy, Vx, j);
    f2v val = {0.0f, 0.0f};
    if (line_ok(g, c, j)) {
      int y0, x0;
      float ty, tx;
      bool yu, xu;
      line_cell(c.y, g.H, y0, ty, yu);
      line_cell(c.x, g.WR, x0, tx, xu);
      const f2v *b = img + (int64_t)y0 * g.WR + x0;
      const f2v zero = {0.0f, 0.0f};
      const f2v v00 = b[0], v01 = xu ? b[1] : zero;
      const f2v v10 = yu ? b[g.WR] : zero, v11 = (yu && xu) ? b[g.WR + 1] : zero;
      const float ux = 1.0f - tx, uy = 1.0f - ty;
      const float r0 = lerp1(v00.x, v01.x, ux, tx), i0 = lerp1(v00.y, v01.y, ux, tx);
      const float r1 = lerp1(v10.x, v11.x, ux, tx), i1 = lerp1(v10.y, v11.y, ux, tx);
      val.x = lerp1(r0, r1, uy, ty);
      val.y = lerp1(i0, i1, uy, ty);
      if (c.conj) val.y = -val.y;
    }
    out[r * g.wr + j] = val;
  }
}

// Backward.  GI: grad_image += the adjoint splat of kappa(grad_line) (kappa =
// conj where the pixel was read conjugated).  GR: grad_rot[r] (yx, row-major)
// += dL/dR through the coordinates: dL/dc = Re(conj(kappa(g)) dS/dc), zero
// where c <= 0 or c >= N-1 (grid_sample's border-clip gradient);
// dc_y/dR[0][1] = a_y s f and dc_x/dR[1][1] = a_x s f.
template <bool GI, bool GR>
__global__ __launch_bounds__(kLineBlock) void line_bwd_kernel(const f2v *__restrict__ gl,
                                                             const f2v *__restrict__ img, LineGeom g,
                                                             const float *__restrict__ rot,
                                                             f2v *__restrict__ gimg,
                                                             double *__restrict__ grot) {
  const int64_t r = blockIdx.x;
  double Vy, Vx;
  line_rot(g, rot + r * 4, Vy, Vx);
  double acc_y = 0.0, acc_x = 0.0;
  for (int j = threadIdx.x; j < g.wr; j += kLineBlock) {
    const LineCoord c = line_coord(g, Vy, Vx, j);
    if (!line_ok(g, c, j)) continue;
    int y0, x0;
    float ty, tx;
    bool yu, xu;
    line_cell(c.y, g.H, y0, ty, yu);
    line_cell(c.x, g.WR, x0, tx, xu);
    f2v gs = gl[r * g.wr + j];
    if (c.conj) gs.y = -gs.y;
    const float ux = 1.0f - tx, uy = 1.0f - ty;
    const int64_t e = (int64_t)y0 * g.WR + x0;
    if (GI) {
      float *gf = reinterpret_cast<float *>(gimg);
      const float w[4] = {uy * ux, uy * tx, ty * ux, ty * tx};
      const bool ok[4] = {true, xu, yu, yu && xu};
      const int64_t off[4] = {0, 1, g.WR, g.WR + 1};
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (!ok[q]) continue;
        unsafeAtomicAdd(gf + 2 * (e + off[q]), w[q] * gs.x);
        unsafeAtomicAdd(gf + 2 * (e + off[q]) + 1, w[q] * gs.y);
      }
    }
    if (GR) {
      const f2v *b = img + e;
      const f2v zero = {0.0f, 0.0f};
      const f2v v00 = b[0], v01 = xu ? b[1] : zero;
      const f2v v10 = yu ? b[g.WR] : zero, v11 = (yu && xu) ? b[g.WR + 1] : zero;
      const float dyr = ux * (v10.x - v00.x) + tx * (v11.x - v01.x);
      const float dyi = ux * (v10.y - v00.y) + tx * (v11.y - v01.y);
      const float dxr = uy * (v01.x - v00.x) + ty * (v11.x - v10.x);
      const float dxi = uy * (v01.y - v00.y) + ty * (v11.y - v10.y);
      float gy = gs.x * dyr + gs.y * dyi;
      float gx = gs.x * dxr + gs.y * dxi;
      if (!(c.y > 0.0 && c.y < (double)(g.H - 1))) gy = 0.0f;
      if (!(c.x > 0.0 && c.x < (double)(g.WR - 1))) gx = 0.0f;
      const double sf = (c.conj ? -1.0 : 1.0) * (double)j * g.inv_W;
      acc_y += (double)gy * g.ay * sf;
      acc_x += (double)gx * g.ax * sf;
    }
  }
  if (GR) {
    __shared__ double red[kLineBlock / 64][2];
    for (int off = 32; off > 0; off >>= 1) {
      acc_y += __shfl_down(acc_y, off, 64);
      acc_x += __shfl_down(acc_x, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
      red[wid][0] = acc_y;
      red[wid][1] = acc_x;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
      double s = 0.0;
      for (int w = 0; w < kLineBlock / 64; ++w) s += red[w][threadIdx.x];
      // one workgroup per rotation: a plain read-modify-write accumulates
      grot[r * 4 + (threadIdx.x ? 3 : 1)] += s;
    }
  }
}

int check_line_args(int h, int wr, int img_h, int img_w, int64_t n_rot, const float *rot) {
  if (h < 1 || wr < 1) return fail_invalid("image_rfft dimensions must be >= 1");
  if (img_h < 2 || img_w < 1) return fail_invalid("image_shape must be (h >= 2, w >= 1)");
  if (n_rot < 0 || n_rot > 0x7fffffffLL) return fail_invalid("rotation count out of range");
  if (n_rot > 0 && !rot) return fail_invalid("rotation pointer is NULL");
  return TT_OK;
}

}  // namespace

}  // namespace tt

using namespace tt;

extern "C" {

int tt_extract2d_c64(const void *image, int h, int wr, int img_h, int img_w, const float *rot_yx,
                     int64_t n_rot, double fftfreq_max, void *lines, void *stream) {
  int st = check_line_args(h, wr, img_h, img_w, n_rot, rot_yx);
  if (st) return st;
  if (n_rot == 0) return TT_OK;
  if (!image || !lines) return fail_invalid("image / lines pointer is NULL");
  const LineGeom g = make_line_geom(h, wr, img_h, img_w, fftfreq_max);
  hipLaunchKernelGGL(line_fwd_kernel, dim3((unsigned)n_rot), dim3(kLineBlock), 0, (hipStream_t)stream,
                     (const f2v *)image, g, rot_yx, (f2v *)lines);
  return check_launch("line_fwd_kernel");
}

int tt_extract2d_bwd_c64(const void *grad_lines, const void *image, int h, int wr, int img_h, int img_w,
                         const float *rot_yx, int64_t n_rot, double fftfreq_max, void *grad_image,
                         double *grad_rot, void *stream) {
  int st = check_line_args(h, wr, img_h, img_w, n_rot, rot_yx);
  if (st) return st;
  if (n_rot == 0 || (!grad_image && !grad_rot)) return TT_OK;
  if (!grad_lines) return fail_invalid("grad_lines pointer is NULL");
  if (grad_rot && !image) return fail_invalid("the image is required for the rotation gradient");
  const LineGeom g = make_line_geom(h, wr, img_h, img_w, fftfreq_max);
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid((unsigned)n_rot), block(kLineBlock);
  if (grad_image && grad_rot)
    hipLaunchKernelGGL((line_bwd_kernel<true, true>), grid, block, 0, s, (const f2v *)grad_lines,
                       (const f2v *)image, g, rot_yx, (f2v *)grad_image, grad_rot);
  else if (grad_image)
    hipLaunchKernelGGL((line_bwd_kernel<true, false>), grid, block, 0, s, (const f2v *)grad_lines,
                       (const f2v *)image, g, rot_yx, (f2v *)grad_image, grad_rot);
  else
    hipLaunchKernelGGL((line_bwd_kernel<false, true>), grid, block, 0, s, (const f2v *)grad_lines,
                       (const f2v *)image, g, rot_yx, (f2v *)grad_image, grad_rot);
  return check_launch("line_bwd_kernel");
}

}  // extern "C"
