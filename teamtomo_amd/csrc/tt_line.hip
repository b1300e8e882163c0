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
// Lines are short (W//2+1 pixels) and read a small image that stays in L2:
// one lane per line pixel over the flat (rotation, pixel) range, no LDS
// staging.  Coordinates are
// evaluated in FP64 and snapped within 1e-9 of an integer, as in the 3-D path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_fixed.h"

namespace tt {

namespace {

constexpr int kLineBlock = 256;

struct LineGeom {
  int H, WR;          // image array: rows, rfft columns
  int wr;             // line length W//2+1 of the image shape
  double ay, by, ax;  // array units per cycle/pixel (y full, x half), y of f = 0
  double inv_W;       // 1/W of the image shape
  double inv_wr;      // 1/wr, for the flat index split
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
  g.inv_wr = 1.0 / (double)g.wr;
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

// Flat (rotation, pixel) index -> rotation, with an FP64 reciprocal and a
// one-step correction (n_rot * wr < 2^53).
__device__ __forceinline__ void split_line(const LineGeom &g, int64_t p, int64_t &r, int &j) {
  r = (int64_t)((double)p * g.inv_wr);
  int64_t jj = p - r * g.wr;
  if (jj < 0) {
    r -= 1;
    jj += g.wr;
  } else if (jj >= g.wr) {
    r += 1;
    jj -= g.wr;
  }
  j = (int)jj;
}

// Bilinear read of the four corners (weight-0 corners past N-1 read as 0).
// The two corners of a row are adjacent: one 16-byte load per row (8-byte
// aligned), which halves the cache-line lookups of a wave whose lanes walk a
// steep line (a new image row per lane).
__device__ __forceinline__ void line_corners(const f2v *__restrict__ img, const LineGeom &g, int y0, int x0,
                                             bool yu, bool xu, f2v (&v)[4]) {
  const f2v *b = img + (int64_t)y0 * g.WR + x0;
  const f2v zero = {0.0f, 0.0f};
  if (xu) {
    const f4a8 r0 = *reinterpret_cast<const f4a8 *>(b);
    v[0] = f2v{r0.x, r0.y};
    v[1] = f2v{r0.z, r0.w};
    if (yu) {
      const f4a8 r1 = *reinterpret_cast<const f4a8 *>(b + g.WR);
      v[2] = f2v{r1.x, r1.y};
      v[3] = f2v{r1.z, r1.w};
    } else {
      v[2] = zero;
      v[3] = zero;
    }
  } else {
    v[0] = b[0];
    v[1] = zero;
    v[2] = yu ? b[g.WR] : zero;
    v[3] = zero;
  }
}

// One lane per line pixel over the flat (rotation, pixel) range; consecutive
// lanes take consecutive pixels of a line, so the corner reads of a wave
// fall on a few image rows that stay in L2.
__global__ __launch_bounds__(kLineBlock) void line_fwd_kernel(const f2v *__restrict__ img, LineGeom g,
                                                             const float *__restrict__ rot, int64_t total,
                                                             f2v *__restrict__ out) {
  for (int64_t p = (int64_t)blockIdx.x * kLineBlock + threadIdx.x; p < total;
       p += (int64_t)gridDim.x * kLineBlock) {
    int64_t r;
    int j;
    split_line(g, p, r, j);
    double Vy, Vx;
    line_rot(g, rot + r * 4, Vy, Vx);
    const LineCoord c = line_coord(g, Vy, Vx, j);
    f2v val = {0.0f, 0.0f};
    if (line_ok(g, c, j)) {
      int y0, x0;
      float ty, tx;
      bool yu, xu;
      line_cell(c.y, g.H, y0, ty, yu);
      line_cell(c.x, g.WR, x0, tx, xu);
      f2v v[4];
      line_corners(img, g, y0, x0, yu, xu, v);
      const float ux = 1.0f - tx, uy = 1.0f - ty;
      const float r0 = lerp1(v[0].x, v[1].x, ux, tx), i0 = lerp1(v[0].y, v[1].y, ux, tx);
      const float r1 = lerp1(v[2].x, v[3].x, ux, tx), i1 = lerp1(v[2].y, v[3].y, ux, tx);
      val.x = lerp1(r0, r1, uy, ty);
      val.y = lerp1(i0, i1, uy, ty);
      if (c.conj) val.y = -val.y;
    }
    out[p] = val;
  }
}

// Backward.  GI: grad_image += the adjoint splat of kappa(grad_line) (kappa =
// conj where the pixel was read conjugated); corners of weight 0 are skipped,
// which keeps the atomics off the DC pixel that every line starts on.
// GR: grad_rot[r] (yx, row-major) += dL/dR through the coordinates:
// dL/dc = Re(conj(kappa(g)) dS/dc), zero where c <= 0 or c >= N-1
// (grid_sample's border-clip gradient); dc_y/dR[0][1] = a_y s f and
// dc_x/dR[1][1] = a_x s f.  Lanes of a wave that share a rotation are summed
// first, then one FP64 atomic per wave and entry.
template <bool GI, bool GR>
__global__ __launch_bounds__(kLineBlock) void line_bwd_kernel(const f2v *__restrict__ gl,
                                                             const f2v *__restrict__ img, LineGeom g,
                                                             const float *__restrict__ rot, int64_t total,
                                                             f2v *__restrict__ gimg,
                                                             double *__restrict__ grot) {
  for (int64_t p0 = (int64_t)blockIdx.x * kLineBlock; p0 < total; p0 += (int64_t)gridDim.x * kLineBlock) {
    const int64_t p = p0 + threadIdx.x;
    double acc_y = 0.0, acc_x = 0.0;
    int64_t r = 0;
    if (p < total) {
      int j;
      split_line(g, p, r, j);
      double Vy, Vx;
      line_rot(g, rot + r * 4, Vy, Vx);
      const LineCoord c = line_coord(g, Vy, Vx, j);
      if (line_ok(g, c, j)) {
        int y0, x0;
        float ty, tx;
        bool yu, xu;
        line_cell(c.y, g.H, y0, ty, yu);
        line_cell(c.x, g.WR, x0, tx, xu);
        f2v gs = gl[p];
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
            if (!ok[q] || w[q] == 0.0f) continue;
            unsafeAtomicAdd(gf + 2 * (e + off[q]), w[q] * gs.x);
            unsafeAtomicAdd(gf + 2 * (e + off[q]) + 1, w[q] * gs.y);
          }
        }
        if (GR) {
          f2v v[4];
          line_corners(img, g, y0, x0, yu, xu, v);
          const float dyr = ux * (v[2].x - v[0].x) + tx * (v[3].x - v[1].x);
          const float dyi = ux * (v[2].y - v[0].y) + tx * (v[3].y - v[1].y);
          const float dxr = uy * (v[1].x - v[0].x) + ty * (v[3].x - v[2].x);
          const float dxi = uy * (v[1].y - v[0].y) + ty * (v[3].y - v[2].y);
          float gy = gs.x * dyr + gs.y * dyi;
          float gx = gs.x * dxr + gs.y * dxi;
          if (!(c.y > 0.0 && c.y < (double)(g.H - 1))) gy = 0.0f;
          if (!(c.x > 0.0 && c.x < (double)(g.WR - 1))) gx = 0.0f;
          const double sf = (c.conj ? -1.0 : 1.0) * (double)j * g.inv_W;
          acc_y = (double)gy * g.ay * sf;
          acc_x = (double)gx * g.ax * sf;
        }
      }
    }
    if (GR) {
      // the wave's lanes cover at most a few rotations: sum each run of
      // lanes with the same rotation (a segmented scan by shuffles), then the
      // last lane of every run adds its sum
      const int lane = threadIdx.x & 63;
      const int64_t rn = __shfl_down(r, 1, 64);
      for (int off = 1; off < 64; off <<= 1) {
        const double uy = __shfl_up(acc_y, off, 64), ux = __shfl_up(acc_x, off, 64);
        const int64_t ru = __shfl_up(r, off, 64);
        if (lane >= off && ru == r) {
          acc_y += uy;
          acc_x += ux;
        }
      }
      const bool last = lane == 63 || rn != r || p + 1 >= total;
      if (p < total && last && (acc_y != 0.0 || acc_x != 0.0)) {
        unsafeAtomicAdd(grot + r * 4 + 1, acc_y);
        unsafeAtomicAdd(grot + r * 4 + 3, acc_x);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Image gradient by tiles: the adjoint splat of the lines, accumulated per
// image tile in LDS.  Global float atomics from every XCD onto the pixels
// around DC (which every line crosses) serialise outside L2, so each unit (an
// image tile and a range of rotations) keeps its tile in LDS: for every
// rotation it intersects the ray with the tile's coordinate box, walks the
// pixels whose cell lies in the tile (the same line_coord / line_cell
// decision as the forward kernel) and adds their bilinear contributions with
// 64-bit LDS atomics in fixed point (scale 2^(30-e) from the exact max|g|;
// 64-bit integer LDS atomics run far faster than float ones); the flush then
// adds the tile to the image, with 64-bit integer global atomics into a
// staging buffer only on the rows and columns a neighbouring tile also
// touches (and for tiles split over several units), converted by one more
// pass: the gradient does not depend on the order of the atomics.
// ---------------------------------------------------------------------------
constexpr int kTY = 32, kTX = 32;  // cells per tile
constexpr int kTP = (kTY + 1) * (kTX + 1);

struct LineUnit {
  int tile;    // ty * ntx + tx
  int r0, r1;  // rotation range
  int split;   // 1 if the tile is shared by several units
};

typedef unsigned long long u64;

// max(|re|, |im|) over n complex values, as float bits, into *res (zeroed
// beforehand); one atomic per block.
__global__ __launch_bounds__(kLineBlock) void line_maxabs_kernel(const f2v *__restrict__ x, int64_t n,
                                                                unsigned *__restrict__ res) {
  __shared__ float s_m[kLineBlock / 64];
  float m = 0.0f;
  // pairs of values as one 16-byte load, four loads in flight per thread
  // (one 8-byte load at a time measured 3.6 TB/s)
  const int64_t npair = n >> 1;
  const f4a8 *x4 = reinterpret_cast<const f4a8 *>(x);
  const int64_t stride = (int64_t)gridDim.x * kLineBlock;
  int64_t i = (int64_t)blockIdx.x * kLineBlock + threadIdx.x;
  for (; i + 3 * stride < npair; i += 4 * stride) {
    f4a8 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = x4[i + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      m = fmaxf(m, fmaxf(fmaxf(fabsf(v[u].x), fabsf(v[u].y)), fmaxf(fabsf(v[u].z), fabsf(v[u].w))));
  }
  for (; i < npair; i += stride) {
    const f4a8 v = x4[i];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {  // the last value of an odd count
    const f2v v = x[n - 1];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = s_m[0];
    for (int w = 1; w < kLineBlock / 64; ++w) b = fmaxf(b, s_m[w]);
    atomicMax(res, __float_as_uint(b));
  }
}

__global__ void line_zero_word_kernel(unsigned *w) {
  if (threadIdx.x == 0) *w = 0u;
}

__global__ __launch_bounds__(kLineBlock) void line_zero_kernel(u64 *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kLineBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kLineBlock)
    p[i] = 0ull;
}

// Fixed-point scale 2^(30 - e), 2^e > max|g|: every contribution w g (|w| <= 1)
// converts to a 32-bit integer, and 2^33 of them fit the 64-bit sums.
__device__ __forceinline__ float line_scale(const unsigned *maxbits) {
  const float m = __uint_as_float(*maxbits);
  int e = 0;
  if (m > 0.0f && m < 3.0e38f) frexpf(m, &e);
  return ldexpf(1.0f, 30 - e);
}

__device__ __forceinline__ int line_cvt_floor(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));  // floor and convert in one VALU op
  return r;
}

__global__ __launch_bounds__(kLineBlock) void line_adj_tile_kernel(const f2v *__restrict__ gl, LineGeom g,
                                                                  const float *__restrict__ rot,
                                                                  const LineUnit *__restrict__ units, int ntx,
                                                                  const unsigned *__restrict__ maxbits,
                                                                  f2v *__restrict__ gimg, u64 *__restrict__ stage) {
  constexpr int NW = kLineBlock / 64;
  // split accumulators (tt_fixed.h): planes of low and high words, re then im
  __shared__ unsigned acc[4][kTP];
  // per wave, the current batch of rays: first candidate, first column,
  // rotation and FP64 line constants, and the chunk-step marks
  __shared__ int s_excl[NW][64], s_j0[NW][64], s_r[NW][64];
  __shared__ double s_vy[NW][64], s_vx[NW][64];
  __shared__ unsigned s_mark[NW][64];
  const LineUnit u = units[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int ty = u.tile / ntx, tx = u.tile - ty * ntx;
  const int y_lo = ty * kTY, x_lo = tx * kTX;
  const int nyc = min(kTY, g.H - 1 - y_lo), nxc = min(kTX, g.WR - 1 - x_lo);  // cells in the tile
  for (int v = tid; v < kTP; v += kLineBlock) {
    acc[0][v] = acc[2][v] = kLoBias;
    acc[1][v] = acc[3][v] = 0u;
  }
  const float vscale = line_scale(maxbits);
  __syncthreads();
  // the tile's coordinate box relative to DC in FP32, widened by 1e-3 (far
  // beyond the FP32 rounding of the box and of the ray, |error| < 2e-4); the
  // candidate columns are widened by 0.01 more, and the exact cell test below
  // decides ownership
  const float yb0 = (float)((double)y_lo - g.by) - 1e-3f, yb1 = (float)((double)(y_lo + nyc) - g.by) + 1e-3f;
  const float xb0 = (float)x_lo - 1e-3f, xb1 = (float)(x_lo + nxc) + 1e-3f;
  const float jmax = (float)(g.wr - 1);
  int *excls = s_excl[wid], *j0s = s_j0[wid], *rs = s_r[wid];
  double *vys = s_vy[wid], *vxs = s_vx[wid];
  unsigned *marks = s_mark[wid];
  // each wave walks its own share of the rotations, skipping batches of 64
  // rotations whose rays all miss the tile without any block barrier
  for (int rb = u.r0 + wid * 64; rb < u.r1; rb += kLineBlock) {
    const int r = rb + lane;
    int cnt = 0, j0 = 0;
    double Vy = 0.0, Vx = 0.0;
    if (r < u.r1) {
      const float R01 = rot[(int64_t)r * 4 + 1], R11 = rot[(int64_t)r * 4 + 3];
      const float vy = (float)(g.ay * g.inv_W) * R01, vx = (float)(g.ax * g.inv_W) * R11;
      // the ray c(j) = (by + s j vy, j |vx|), s = -1 iff vx < 0
      const float sy = vx < 0.0f ? -vy : vy, ax = fabsf(vx);
      float lo = 0.0f, hi = jmax;
      if (fabsf(sy) > 1e-20f) {
        const float rs_ = 1.0f / sy;
        const float t0 = yb0 * rs_, t1 = yb1 * rs_;
        lo = fmaxf(lo, fminf(t0, t1));
        hi = fminf(hi, fmaxf(t0, t1));
      } else if (yb0 > 0.0f || yb1 < 0.0f) {  // c_y = by for every j
        hi = -1.0f;
      }
      if (ax > 1e-20f) {
        const float ra = 1.0f / ax;
        lo = fmaxf(lo, xb0 * ra);
        hi = fminf(hi, xb1 * ra);
      } else if (xb0 > 0.0f) {  // c_x = 0 for every j
        hi = -1.0f;
      }
      if (lo <= hi + 0.02f) {
        j0 = max(0, (int)ceilf(lo - 0.01f));
        const int j1 = min(g.wr - 1, (int)floorf(hi + 0.01f));
        cnt = max(0, j1 - j0 + 1);
      }
      line_rot(g, rot + (int64_t)r * 4, Vy, Vx);
    }
    const int incl = wave_incl_scan(cnt);
    const int excl = incl - cnt;
    const int total = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(incl, 63));
    if (total == 0) continue;  // wave-uniform
    excls[lane] = excl;
    j0s[lane] = j0;
    rs[lane] = r;
    vys[lane] = Vy;
    vxs[lane] = Vx;
    // candidates 64 at a time: the rays whose candidates start inside the
    // step mark their lane (+1) there, and a running maximum over the wave
    // carries each mark to the candidates that follow
    for (int qb = 0; qb < total; qb += 64) {
      const unsigned long long hb = __ballot(excl <= qb && qb < incl);
      const int k0 = (int)__builtin_ctzll(hb);
      __builtin_amdgcn_wave_barrier();  // the previous step's marks have been read
      marks[lane] = lane == 0 ? (unsigned)(k0 + 1) : 0u;
      if (cnt > 0 && excl > qb && excl < qb + 64) marks[excl - qb] = (unsigned)(lane + 1);
      __builtin_amdgcn_wave_barrier();
      const int k = (int)wave_incl_max(marks[lane]) - 1;
      const int q = qb + lane;
      if (q >= total) continue;
      const int j = j0s[k] + (q - excls[k]);
      const int rr = rs[k];
      const LineCoord c = line_coord(g, vys[k], vxs[k], j);
      if (!line_ok(g, c, j)) continue;
      int y0, x0;
      float fy, fx;
      bool yu, xu;
      line_cell(c.y, g.H, y0, fy, yu);
      line_cell(c.x, g.WR, x0, fx, xu);
      const int ly = y0 - y_lo, lx = x0 - x_lo;
      if (ly < 0 || ly >= nyc || lx < 0 || lx >= nxc) continue;  // another tile's cell
      f2v gs = gl[(int64_t)rr * g.wr + j];
      if (c.conj) gs.y = -gs.y;
      gs.x *= vscale;
      gs.y *= vscale;
      const float ux = 1.0f - fx, uy = 1.0f - fy;
      const float w[4] = {uy * ux, uy * fx, fy * ux, fy * fx};
      const int o[4] = {0, 1, kTX + 1, kTX + 2};
      const int e = ly * (kTX + 1) + lx;
      // round(w g) (half up) added to the low words; the rare carries below
      unsigned old[4][2];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        old[t][0] = atomicAdd(&acc[0][e + o[t]], (unsigned)line_cvt_floor(__fmaf_rn(w[t], gs.x, 0.5f)));
        old[t][1] = atomicAdd(&acc[2][e + o[t]], (unsigned)line_cvt_floor(__fmaf_rn(w[t], gs.y, 0.5f)));
      }
      if (!lo_all_in_window<8>(&old[0][0])) {
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const unsigned h0 = hi_step(old[t][0], line_cvt_floor(__fmaf_rn(w[t], gs.x, 0.5f)));
          const unsigned h1 = hi_step(old[t][1], line_cvt_floor(__fmaf_rn(w[t], gs.y, 0.5f)));
          if (h0) atomicAdd(&acc[1][e + o[t]], h0);
          if (h1) atomicAdd(&acc[3][e + o[t]], h1);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();  // the next batch rewrites this wave's arrays
  }
  __syncthreads();
  // flush: exclusive pixels by a plain read-modify-write, shared ones (the
  // tile's first row / column, its far halo row / column, or all of a split
  // tile) as 64-bit integer atomics into the staging sums, which
  // line_stage_flush_kernel converts: integer sums do not depend on the
  // order of the atomics, so the gradient is bitwise reproducible
  const float inv = 1.0f / vscale;
  for (int v = tid; v < kTP; v += kLineBlock) {
    const int ly = v / (kTX + 1), lx = v - ly * (kTX + 1);
    if (ly > nyc || lx > nxc) continue;
    const u64 qr = split_value(acc[0][v], acc[1][v]), qi = split_value(acc[2][v], acc[3][v]);
    if (qr == 0ull && qi == 0ull) continue;
    const float ar = (float)(long long)qr * inv, ai = (float)(long long)qi * inv;
    const int64_t e = (int64_t)(y_lo + ly) * g.WR + (x_lo + lx);
    const bool shared = u.split || ly == 0 || lx == 0 || ly == nyc || lx == nxc;
    if (shared) {
      atomicAdd(&stage[2 * e], qr);
      atomicAdd(&stage[2 * e + 1], qi);
    } else {
      f2v cur = gimg[e];
      cur.x += ar;
      cur.y += ai;
      gimg[e] = cur;
    }
  }
}

// Adds the staged integer sums of the shared pixels to the image gradient,
// at the scale of the tile kernel.
__global__ __launch_bounds__(kLineBlock) void line_stage_flush_kernel(const u64 *__restrict__ stage, int64_t n_px,
                                                                     const unsigned *__restrict__ maxbits,
                                                                     f2v *__restrict__ gimg) {
  const float inv = 1.0f / line_scale(maxbits);
  for (int64_t e = (int64_t)blockIdx.x * kLineBlock + threadIdx.x; e < n_px; e += (int64_t)gridDim.x * kLineBlock) {
    const u64 qr = stage[2 * e], qi = stage[2 * e + 1];
    if (qr == 0ull && qi == 0ull) continue;
    f2v cur = gimg[e];
    cur.x += (float)(long long)qr * inv;
    cur.y += (float)(long long)qi * inv;
    gimg[e] = cur;
  }
}

// Units of the tiles: a tile's work is estimated for uniformly spread angles
// (the fraction of rays through DC that cross the tile is the angle it
// subtends over pi) and heavy tiles are split into rotation ranges.
struct LinePlanKey {
  int dev, H, WR, wr, nrot;
  bool operator<(const LinePlanKey &o) const {
    return std::tie(dev, H, WR, wr, nrot) < std::tie(o.dev, o.H, o.WR, o.wr, o.nrot);
  }
};
std::mutex g_line_mutex;
std::map<LinePlanKey, std::pair<LineUnit *, int>> g_line_plans;

int line_plan(const LineGeom &g, int64_t n_rot, LineUnit **units, int *n_units, int *ntx_out) {
  int dev = 0;
  if (hipError_t e = hipGetDevice(&dev)) {
    set_error(std::string("line plan: hipGetDevice: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  const int nty = (g.H - 1 + kTY - 1) / kTY, ntx = (g.WR - 1 + kTX - 1) / kTX;
  *ntx_out = ntx;
  const LinePlanKey key{dev, g.H, g.WR, g.wr, (int)n_rot};
  std::lock_guard<std::mutex> lock(g_line_mutex);
  auto it = g_line_plans.find(key);
  if (it != g_line_plans.end()) {
    *units = it->second.first;
    *n_units = it->second.second;
    return TT_OK;
  }
  const double kPi = 3.14159265358979323846;
  std::vector<double> est(nty * ntx);
  double total = 0.0;
  for (int ty = 0; ty < nty; ++ty)
    for (int tx = 0; tx < ntx; ++tx) {
      // the tile's box relative to DC (by, 0); x >= 0 throughout
      const double y0 = ty * kTY - g.by, y1 = std::min((ty + 1) * kTY, g.H - 1) - g.by;
      const double x0 = tx * kTX, x1 = std::min((tx + 1) * kTX, g.WR - 1);
      double frac = 1.0;
      const double dist = std::hypot(x0, (y0 <= 0.0 && y1 >= 0.0) ? 0.0 : std::min(std::fabs(y0), std::fabs(y1)));
      if (dist > 2.0) {
        double amin = 1e9, amax = -1e9;
        const double ys[2] = {y0, y1}, xs[2] = {x0, x1};
        for (double yy : ys)
          for (double xx : xs) {
            const double a = std::atan2(xx, yy);  // in [0, pi]
            amin = std::min(amin, a);
            amax = std::max(amax, a);
          }
        // subtended angle plus the rays' one-pixel bilinear reach
        frac = std::min(1.0, (amax - amin + 2.0 / dist) / kPi);
      }
      // per rotation a range test; per crossing ray about a tile width of pixels
      est[ty * ntx + tx] = (double)n_rot * (0.05 + frac * kTX);
      total += est[ty * ntx + tx];
    }
  // about 8 units per CU of 256, none below a few range tests per rotation
  const double target = std::max(total / (256.0 * 8.0), (double)n_rot * 0.2);
  struct U {
    LineUnit u;
    double w;
  };
  std::vector<U> hu;
  for (int t = 0; t < nty * ntx; ++t) {
    const int nseg = (int)std::max<int64_t>(
        1, std::min<int64_t>(n_rot / 256 + 1, (int64_t)std::ceil(est[t] / target)));
    for (int sgi = 0; sgi < nseg; ++sgi) {
      LineUnit u;
      u.tile = t;
      u.r0 = (int)(n_rot * sgi / nseg);
      u.r1 = (int)(n_rot * (sgi + 1) / nseg);
      u.split = nseg > 1;
      if (u.r1 > u.r0) hu.push_back({u, est[t] / nseg});
    }
  }
  // heaviest units first
  std::stable_sort(hu.begin(), hu.end(), [](const U &a, const U &b) { return a.w > b.w; });
  std::vector<LineUnit> h(hu.size());
  for (size_t i = 0; i < hu.size(); ++i) h[i] = hu[i].u;
  LineUnit *d = nullptr;
  if (hipError_t e = hipMalloc(&d, h.size() * sizeof(LineUnit))) {
    set_error(std::string("line plan: hipMalloc: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  if (hipError_t e = hipMemcpy(d, h.data(), h.size() * sizeof(LineUnit), hipMemcpyHostToDevice)) {
    set_error(std::string("line plan: hipMemcpy: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  g_line_plans[key] = {d, (int)h.size()};
  *units = d;
  *n_units = (int)h.size();
  return TT_OK;
}

int line_grid(int64_t total) {
  const int64_t b = (total + kLineBlock - 1) / kLineBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 1 << 16));
}

int check_line_args(int h, int wr, int img_h, int img_w, int64_t n_rot, const float *rot) {
  if (h < 1 || wr < 1) return fail_invalid("image_rfft dimensions must be >= 1");
  if (img_h < 2 || img_w < 1) return fail_invalid("image_shape must be (h >= 2, w >= 1)");
  if (n_rot < 0 || n_rot > ((int64_t)1 << 40)) return fail_invalid("rotation count out of range");
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
  const int64_t total = n_rot * (int64_t)g.wr;
  hipLaunchKernelGGL(line_fwd_kernel, dim3(line_grid(total)), dim3(kLineBlock), 0, (hipStream_t)stream,
                     (const f2v *)image, g, rot_yx, total, (f2v *)lines);
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
  const int64_t total = n_rot * (int64_t)g.wr;
  const dim3 grid(line_grid(total)), block(kLineBlock);
  // the image gradient by tiles (images of at least 2 x 2 samples and rotation
  // counts the plan indexes with int), else by global atomics
  const bool tiles = grad_image && h >= 2 && wr >= 2 && n_rot < ((int64_t)1 << 30);
  if (tiles) {
    LineUnit *units = nullptr;
    int n_units = 0, ntx = 0;
    if ((st = line_plan(g, n_rot, &units, &n_units, &ntx))) return st;
    // scratch: the scale word, then two 64-bit staging sums per image pixel
    const int64_t n_px = (int64_t)h * wr;
    void *scratch = nullptr;
    if (hipError_t e = hipMallocAsync(&scratch, 256 + (size_t)n_px * 2 * sizeof(u64), s)) {
      set_error(std::string("tt_extract2d_bwd_c64: hipMallocAsync: ") + hipGetErrorString(e));
      return TT_ERR_HIP;
    }
    unsigned *word = (unsigned *)scratch;
    u64 *stage = (u64 *)((char *)scratch + 256);
    const unsigned zgrid = (unsigned)std::min<int64_t>(4096, (2 * n_px + kLineBlock - 1) / kLineBlock);
    hipLaunchKernelGGL(line_zero_word_kernel, dim3(1), dim3(64), 0, s, word);
    hipLaunchKernelGGL(line_zero_kernel, dim3(zgrid), block, 0, s, stage, 2 * n_px);
    hipLaunchKernelGGL(line_maxabs_kernel, dim3((unsigned)std::min<int64_t>(2048, (total + kLineBlock - 1) / kLineBlock)),
                       block, 0, s, (const f2v *)grad_lines, total, word);
    hipLaunchKernelGGL(line_adj_tile_kernel, dim3((unsigned)n_units), block, 0, s, (const f2v *)grad_lines,
                       g, rot_yx, (const LineUnit *)units, ntx, (const unsigned *)word, (f2v *)grad_image, stage);
    hipLaunchKernelGGL(line_stage_flush_kernel, dim3((unsigned)std::min<int64_t>(4096, (n_px + kLineBlock - 1) / kLineBlock)),
                       block, 0, s, (const u64 *)stage, n_px, (const unsigned *)word, (f2v *)grad_image);
    st = check_launch("line_adj_tile_kernel / line_stage_flush_kernel");
    if (hipError_t e = hipFreeAsync(scratch, s)) {
      if (!st) {
        set_error(std::string("tt_extract2d_bwd_c64: hipFreeAsync: ") + hipGetErrorString(e));
        st = TT_ERR_HIP;
      }
    }
    if (st) return st;
  }
  if (grad_image && grad_rot && !tiles)
    hipLaunchKernelGGL((line_bwd_kernel<true, true>), grid, block, 0, s, (const f2v *)grad_lines,
                       (const f2v *)image, g, rot_yx, total, (f2v *)grad_image, grad_rot);
  else if (grad_image && !tiles)
    hipLaunchKernelGGL((line_bwd_kernel<true, false>), grid, block, 0, s, (const f2v *)grad_lines,
                       (const f2v *)image, g, rot_yx, total, (f2v *)grad_image, grad_rot);
  else if (grad_rot)
    hipLaunchKernelGGL((line_bwd_kernel<false, true>), grid, block, 0, s, (const f2v *)grad_lines,
                       (const f2v *)image, g, rot_yx, total, (f2v *)grad_image, grad_rot);
  return check_launch("line_bwd_kernel");
}

}  // extern "C"
