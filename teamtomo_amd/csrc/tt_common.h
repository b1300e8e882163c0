// Shared device/host helpers for the gfx950 Fourier-slice kernels.
//
// Geometry conventions follow the reference (packages/primitives/
// torch-fourier-slice/src/torch_fourier_slice): slice pixel (i, j) has
// frequency (0, fy, fx) with fy = (i - H//2)/H, fx = j/W
// (_grids/_central_slice_fftfreq_grid.py:6-32); rotated frequencies map to
// array coordinates c = (f + 0.5)(N-1)/(1-1/N) on full axes and
// c = 2 f (Wr-1) on the half axis (_dft_utils.py:59-67).  Coordinates are
// evaluated in FP64 (full-rate on CDNA4 vector ALUs) and snapped to integers
// within 1e-9 voxel, so exact-geometry rotations land exactly on voxels and
// faces as the reference's float32 arithmetic does.
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
  float invWr_f;          // 1/Wr for the flat-index split
};

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

// Rotation loaded once per block (wave-uniform -> scalar registers).
struct Rot {
  double r01, r02, r11, r12, r21, r22;  // columns y, x of the zyx matrix
  bool tie;                             // extraction conj decision on the x'=0 plane
};

__device__ __forceinline__ Rot load_rot(const float *__restrict__ R, const SliceGeom &g) {
  Rot r;
  r.r01 = (double)R[1];
  r.r02 = (double)R[2];
  r.r11 = (double)R[4];
  r.r12 = (double)R[5];
  r.r21 = (double)R[7];
  r.r22 = (double)R[8];
  // _determine_near_zero_conjugate_mask (:53-79): the sign of R_zyx[2,2]
  // decides, or that of R_zyx[2,1] when |R_zyx[2,2]| <= 10 tol.
  r.tie = (fabs(r.r22) <= g.xtol10) ? (r.r21 < 0.0) : (r.r22 < 0.0);
  return r;
}

// Rotated, Friedel-folded array coordinate of slice pixel (i, j).
// INSERT: conjugate iff x' < 0 (insertion :78-80).  Otherwise the extraction
// rule: x' < -tol, or |x'| < tol and the per-rotation tie (extraction :193-203).
template <bool INSERT>
__device__ __forceinline__ Coord slice_coord(const SliceGeom &g, const Rot &r, int i, int j) {
  double fy = (double)(i - g.half_H) * g.inv_H;
  double fx = (double)j * g.inv_W;
  double qz = fma(r.r01, fy, r.r02 * fx);
  double qy = fma(r.r11, fy, r.r12 * fx);
  double qx = fma(r.r21, fy, r.r22 * fx);
  Coord c;
  if (INSERT)
    c.conj = qx < 0.0;
  else
    c.conj = (qx < -g.xtol) || (fabs(qx) < g.xtol && r.tie);
  double s = c.conj ? -1.0 : 1.0;
  c.z = snap(fma(s * qz, g.az, g.bz));
  c.y = snap(fma(s * qy, g.ay, g.by));
  c.x = snap(s * qx * g.ax);
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

// Host-side geometry setup shared by extraction and insertion.
SliceGeom make_geom(int img_d, int img_h, int img_w, int vol_d, int vol_h, int vol_wr,
                    double fftfreq_max, double x_tolerance);

// Thread-local error message handling for the C ABI.
void set_error(const std::string &msg);
int fail_invalid(const std::string &msg);
int check_launch(const char *what);

}  // namespace tt
