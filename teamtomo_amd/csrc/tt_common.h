// Shared device/host helpers for the gfx950 Fourier-slice kernels.
//
// Geometry conventions follow the reference (packages/primitives/
// torch-fourier-slice/src/torch_fourier_slice): slice pixel (i, j) has
// frequency (0, fy, fx) with fy = (i - H//2)/H, fx = j/W
// (_grids/_central_slice_fftfreq_grid.py:6-32); rotated frequencies map to
// array coordinates c = (f + 0.5)(N-1)/(1-1/N) on full axes and
// c = 2 f (Wr-1) on the half axis (_dft_utils.py:59-67).
//
// Canonical coordinates.  Every kernel (pixel gather, brick gather and
// scatter, zero pass, backward passes) evaluates the coordinate of pixel
// (i, j) under rotation R and Friedel sign s with the same FP64 operations:
//   U_d = a_d R[d][1],  V_d = (a_d R[d][2]) / W,  fy = (i - H//2) / H
//   c_d = fma(j, s V_d, fma(s U_d, fy, b_d)),   b = (bz, by, 0)
// and snaps values within 1e-9 of an integer to that integer, so that
// exact-geometry rotations land exactly on voxels and faces as the
// reference's float32 arithmetic does, and so that a brick kernel and the
// pixel kernels agree bit for bit on which cell (and brick) owns a pixel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace tt {

constexpr double kSnapEps = 1e-9;

typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4a8 __attribute__((ext_vector_type(4), aligned(8)));
typedef double d2v __attribute__((ext_vector_type(2)));

// Complex element type per precision.
template <typename T> struct Cx;
template <> struct Cx<float> { typedef f2v type; };
template <> struct Cx<double> { typedef d2v type; };

// Geometry of one slice stack against one volume, passed by value.
struct SliceGeom {
  int H, Wr, half_H, P;   // slice rows, rfft columns, H/2, H*Wr
  double inv_H, inv_W;    // 1/H, 1/W of the image shape
  double az, ay, ax;      // array units per cycle/pixel (z, y full; x half)
  double bz, by;          // 0.5*az, 0.5*ay: array coordinate of f = 0
  int vD, vH, vW;         // volume shape (vW = rfft width)
  int use_mask;           // fftfreq_max given
  float fmax;             // float32(fftfreq_max)
  float inv_H32, inv_W32; // float32(1/H), float32(1/W) as torch.fft.fftfreq uses
  double xtol, xtol10;    // float32(x_tol), float32(10 x_tol) as doubles
  double tolx;            // xtol * ax: the x' tolerance in half-axis array units
  float invWr_f;          // 1/Wr for the flat-index split
  int64_t ostride;        // gather output elements per rotation (P; channels x P when multichannel)
  int64_t vstride;        // input volume elements per channel (vD vH vW)
};

// Inclusive prefix sum over the wave with DPP lane moves (row_shr 1, 2, 4, 8
// inside rows of 16 lanes, then row_bcast 15 and 31 across rows): six VALU
// operations and no LDS traffic.  Call with every lane active.
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return v;
}

// Inclusive running maximum of unsigned values over the wave (same lane moves).
__device__ __forceinline__ unsigned wave_incl_max(unsigned v) {
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
  v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
  return v;
}

// Flat pixel index -> (row, column) without an integer division.
__device__ __forceinline__ void split_pixel(const SliceGeom &g, int p, int &i, int &j) {
  i = (int)(((float)p + 0.5f) * g.invWr_f);
  j = p - i * g.Wr;
  if (j < 0) {
    i -= 1;
    j += g.Wr;
  } else if (j >= g.Wr) {
    i += 1;
    j -= g.Wr;
  }
}

// The fftfreq_max disc, evaluated in float32 exactly as the reference does
// (slice_extraction/_extract_central_slices_rfft_3d.py:170-180): torch's
// float32 fftfreq values, squares and sum rounded separately, IEEE sqrt.
__device__ __forceinline__ bool in_disc(const SliceGeom &g, int i, int j) {
  if (!g.use_mask) return true;
  float fy = __fmul_rn((float)(i - g.half_H), g.inv_H32);
  float fx = __fmul_rn((float)j, g.inv_W32);
  float r2 = __fadd_rn(__fmul_rn(fy, fy), __fmul_rn(fx, fx));
  return __fsqrt_rn(r2) <= g.fmax;
}

__device__ __forceinline__ double snap(double c) {
  double r = rint(c);
  return fabs(c - r) < kSnapEps ? r : c;
}

struct Coord {
  double z, y, x;
  bool conj;
};

// Canonical per-rotation constants (wave-uniform when one rotation per block).
struct Rot {
  double U[3], V[3];  // a_d R_zyx[d][1] and (a_d R_zyx[d][2]) / W, d = z, y, x
  bool tie;           // extraction conj decision on the x'=0 plane
};

__device__ __forceinline__ Rot load_rot(const float *__restrict__ R, const SliceGeom &g) {
  Rot r;
  const double a[3] = {g.az, g.ay, g.ax};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.U[d] = __dmul_rn(a[d], (double)R[3 * d + 1]);
    r.V[d] = __dmul_rn(__dmul_rn(a[d], (double)R[3 * d + 2]), g.inv_W);
  }
  // _determine_near_zero_conjugate_mask (:53-79): the sign of R_zyx[2,2]
  // decides, or that of R_zyx[2,1] when |R_zyx[2,2]| <= 10 tol.
  const double r21 = (double)R[7], r22 = (double)R[8];
  r.tie = (fabs(r22) <= g.xtol10) ? (r21 < 0.0) : (r22 < 0.0);
  return r;
}

__device__ __forceinline__ double canon_fy(const SliceGeom &g, int i) {
  return __dmul_rn((double)(i - g.half_H), g.inv_H);
}

// x coordinate for s = +1 (a_x x'), whose sign decides the Friedel half.
__device__ __forceinline__ double canon_cx1(const Rot &r, double fy, int j) {
  return __fma_rn((double)j, r.V[2], __dmul_rn(r.U[2], fy));
}

// INSERT: conjugate iff x' < 0 (insertion :78-80).  Otherwise the extraction
// rule: x' < -tol, or |x'| < tol and the per-rotation tie (extraction :193-203).
template <bool INSERT>
__device__ __forceinline__ bool canon_conj(const SliceGeom &g, bool tie, double cx1) {
  if (INSERT) return cx1 < 0.0;
  return (cx1 < -g.tolx) || (fabs(cx1) < g.tolx && tie);
}

// Rotated, Friedel-folded, snapped array coordinate of slice pixel (i, j).
template <bool INSERT>
__device__ __forceinline__ Coord slice_coord(const SliceGeom &g, const Rot &r, int i, int j) {
  const double fy = canon_fy(g, i);
  const double jd = (double)j;
  const double cx1 = canon_cx1(r, fy, j);
  Coord c;
  c.conj = canon_conj<INSERT>(g, r.tie, cx1);
  const double s = c.conj ? -1.0 : 1.0;
  c.z = snap(__fma_rn(jd, s * r.V[0], __fma_rn(s * r.U[0], fy, g.bz)));
  c.y = snap(__fma_rn(jd, s * r.V[1], __fma_rn(s * r.U[1], fy, g.by)));
  c.x = snap(c.conj ? -cx1 : cx1);
  return c;
}

__device__ __forceinline__ bool inside(const SliceGeom &g, const Coord &c) {
  return c.z >= 0.0 && c.z <= (double)(g.vD - 1) && c.y >= 0.0 && c.y <= (double)(g.vH - 1) &&
         c.x >= 0.0 && c.x <= (double)(g.vW - 1);
}

// Lower corner clamped to N-2 so the upper corner is always addressable; the
// fraction is then 1 at c = N-1, which gives the same trilinear weights as
// dropping the out-of-range corner.
struct Cell {
  int z0, y0, x0;
  float tz, ty, tx;
};

__device__ __forceinline__ Cell cell_of(const SliceGeom &g, const Coord &c) {
  Cell k;
  k.z0 = min((int)c.z, g.vD - 2);
  k.y0 = min((int)c.y, g.vH - 2);
  k.x0 = min((int)c.x, g.vW - 2);
  k.tz = (float)(c.z - (double)k.z0);
  k.ty = (float)(c.y - (double)k.y0);
  k.tx = (float)(c.x - (double)k.x0);
  return k;
}

// Trilinear interpolation of 8 complex corners, x then y then z, with the
// products and fused adds pinned so every gather kernel rounds alike.
__device__ __forceinline__ float lerp1(float a, float b, float u, float t) {
  return __fmaf_rn(b, t, __fmul_rn(a, u));
}

__device__ __forceinline__ f2v trilerp(f2v v000, f2v v001, f2v v010, f2v v011, f2v v100, f2v v101,
                                       f2v v110, f2v v111, float tz, float ty, float tx) {
  const float ux = 1.0f - tx, uy = 1.0f - ty, uz = 1.0f - tz;
  const float r00 = lerp1(v000.x, v001.x, ux, tx), i00 = lerp1(v000.y, v001.y, ux, tx);
  const float r01 = lerp1(v010.x, v011.x, ux, tx), i01 = lerp1(v010.y, v011.y, ux, tx);
  const float r10 = lerp1(v100.x, v101.x, ux, tx), i10 = lerp1(v100.y, v101.y, ux, tx);
  const float r11 = lerp1(v110.x, v111.x, ux, tx), i11 = lerp1(v110.y, v111.y, ux, tx);
  const float r0 = lerp1(r00, r01, uy, ty), i0 = lerp1(i00, i01, uy, ty);
  const float r1 = lerp1(r10, r11, uy, ty), i1 = lerp1(i10, i11, uy, ty);
  f2v out;
  out.x = lerp1(r0, r1, uz, tz);
  out.y = lerp1(i0, i1, uz, tz);
  return out;
}

// Host-side geometry setup shared by extraction and insertion.
SliceGeom make_geom(int img_d, int img_h, int img_w, int vol_d, int vol_h, int vol_wr,
                    double fftfreq_max, double x_tolerance);

// Thread-local error message handling for the C ABI.
void set_error(const std::string &msg);
int fail_invalid(const std::string &msg);
int check_launch(const char *what);

}  // namespace tt
