// Brick-stationary scatter for gfx950: the extraction adjoint and the
// central-slice insertion, without global float atomics.
//
// Each workgroup owns one BZ x BY x BX brick of the output volume and keeps
// its accumulator in LDS as 64-bit fixed point: every contribution w * v is
// scaled by 2^k (k from max|v|, so |w v 2^k| <= 2^30), rounded to an integer
// and added with ds_add_u64.  Integer adds commute, so the result is bitwise
// reproducible run to run, and integer LDS atomics run at close to LDS
// bandwidth on CDNA4 whereas ds_add_f32 is many times slower.
//
// Work of one block:
//  (1) one lane per rotation tests the slice plane against the brick's
//      footprint box and, for each crossing (rotation, Friedel half), records
//      the range of slice rows whose pixels can land in the box;
//  (2) the items' rows are pooled: one lane per row (rows of several items
//      share a wave) computes the row's pixel interval from three slab
//      constraints in brick-local FP32 coordinates, taken relative to the
//      item's pixel nearest the brick so every offset stays small; rows are
//      cut into chunks of 4 pixels and a wave prefix sum hands one chunk to
//      each lane;
//  (3) a lane walks its chunk (c = c_row + q * s * B); pixels within 1e-3
//      voxel of a volume face or of the x'=0 Friedel plane take the exact
//      FP64 path (the snapped coordinates of the gather), the rest go
//      straight to the 8-corner splat;
//  (4) the brick is written once.  Bricks in the DC zone, which every slice
//      crosses, are split over rotation segments; their partial sums meet in
//      a 64-bit global scratch and a combine pass writes them.
//
// Insertion semantics (packages/primitives/torch-fourier-slice/src/
// torch_fourier_slice/slice_insertion/_insert_central_slices_rfft_3d.py):
// the direct splat goes to the bricks; the x=0-plane Friedel mirror (:348-404)
// is applied afterwards from a copy of the plane:
//   V[z,y,0] = P[z,y] + conj(P[mz,my])   unless (mz,my) is DC,
// with (mz,my) = ((D-z)%D, (H-y)%H), because the mirror of every x=0 corner
// other than DC is added conjugated at that point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_scatter.h"

namespace tt {

namespace {

constexpr int kT = 256;            // threads of the helper kernels
constexpr int kRotChunk = 64;      // rotations tested per pass, one lane each
constexpr int kItemsMax = 2 * kRotChunk;
constexpr double kMargin = 1e-6;   // slack of the row-interval bounds
constexpr float kGuard = 1e-3f;    // FP32 guard band (voxels) -> exact FP64 path
constexpr float kWeightScale = 1073741824.0f;  // 2^30

typedef unsigned long long u64;

// LDS layout of a brick accumulator with a one-voxel margin on every side:
// local coordinates run from -1 to B.  Corners of a cell that straddles the
// brick face land in the margin, which is never flushed, so every pixel adds
// all 8 corners without per-corner tests (no divergence between lanes).
template <int BZ, int BY, int BX> struct Lay {
  static constexpr int PY = BX + 2;
  static constexpr int PZ = (BY + 2) * PY;
  static constexpr int SIZE = (BZ + 2) * PZ;
  __device__ static int at(int z, int y, int x) { return (z + 1) * PZ + (y + 1) * PY + (x + 1); }
};

// Brick grid; bricks of the DC zone are split into rotation segments.
struct BrickGrid {
  int nbz, nby, nbx, n_bricks;
  int zb0, zb1, yb0, yb1, xb1;  // zone: bz in [zb0,zb1], by in [yb0,yb1], bx in [0,xb1]
  int nzy, nzx, n_zone;         // zone extents (bricks)
  int nseg;                     // segments per zone brick
  int n_units, n_units_padded;
};

struct PlaneGeom {
  double A[3], B[3];
};

__device__ __forceinline__ PlaneGeom plane_geom(const SliceGeom &g, const float *R) {
  PlaneGeom p;
  const double a[3] = {g.az, g.ay, g.ax};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    p.A[d] = a[d] * (double)R[3 * d + 1] * g.inv_H;
    p.B[d] = a[d] * (double)R[3 * d + 2] * g.inv_W;
  }
  return p;
}

// Per-rotation plane data shared by every brick: the slice axes A (rows)
// and B (columns) in array units, the plane normal, and the rows of the
// pseudo-inverse that map a point of the plane back to (row, column).
struct PlanePre {
  double A[3], B[3], n[3], pm[3], pj[3];
  double nn;  // |n|; negative for a degenerate matrix (every pixel a candidate)
};

__global__ __launch_bounds__(kT) void plane_prep_kernel(SliceGeom g, const float *__restrict__ rot,
                                                        int n_rot, PlanePre *__restrict__ pre) {
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= n_rot) return;
  const PlaneGeom p = plane_geom(g, rot + (int64_t)r * 9);
  PlanePre o;
  for (int d = 0; d < 3; ++d) {
    o.A[d] = p.A[d];
    o.B[d] = p.B[d];
  }
  o.n[0] = p.A[1] * p.B[2] - p.A[2] * p.B[1];
  o.n[1] = p.A[2] * p.B[0] - p.A[0] * p.B[2];
  o.n[2] = p.A[0] * p.B[1] - p.A[1] * p.B[0];
  o.nn = sqrt(o.n[0] * o.n[0] + o.n[1] * o.n[1] + o.n[2] * o.n[2]);
  const double AA = p.A[0] * p.A[0] + p.A[1] * p.A[1] + p.A[2] * p.A[2];
  const double BB = p.B[0] * p.B[0] + p.B[1] * p.B[1] + p.B[2] * p.B[2];
  const double AB = p.A[0] * p.B[0] + p.A[1] * p.B[1] + p.A[2] * p.B[2];
  const double det = AA * BB - AB * AB;
  if (det > 1e-20 * AA * BB) {
    for (int d = 0; d < 3; ++d) {
      o.pm[d] = (BB * p.A[d] - AB * p.B[d]) / det;
      o.pj[d] = (AA * p.B[d] - AB * p.A[d]) / det;
    }
  } else {
    for (int d = 0; d < 3; ++d) o.pm[d] = o.pj[d] = 0.0;
    o.nn = -1.0;
  }
  pre[r] = o;
}

// Adds round-toward-zero fixed point x; |x| <= 2^30 by the choice of scale.
__device__ __forceinline__ void lds_add(u64 *p, float x) {
  atomicAdd(p, (u64)(long long)(int)x);
}

// Fixed-point value scale 2^(30-e) with max|v| < 2^e, so |w v scale| <= 2^30.
__device__ __forceinline__ float value_scale(const float *maxabs) {
  const float m = *maxabs;
  if (!(m > 0.0f) || !(m < 3.0e38f)) return 1.0f;
  int e;
  frexpf(m, &e);
  return ldexpf(1.0f, 30 - e);
}

// Threads per brick block: the adjoint's LDS (u64 re/im accumulators, items,
// row tables) allows two 512-thread blocks per CU; the insertion's extra
// weight accumulator allows two 256-thread blocks.
template <bool INSERT> struct ScatterCfg {
  static constexpr int T = INSERT ? 256 : 512;
  static constexpr int W = T / 64;
  static constexpr int MIN_WAVES_PER_SIMD = INSERT ? 2 : 4;  // two blocks per CU
};

// One slice row's candidate pixels in brick-local FP32 coordinates:
// pixel q of the row (column jl + q) sits at st + q * sb.
struct RowRec {
  int jl, excl, i, code;  // first column, pixels before this row, slice row, item code
  float st0, st1, st2;
  float sb0, sb1, sb2;
};

template <int BZ, int BY, int BX, bool INSERT>
__global__ __launch_bounds__(ScatterCfg<INSERT>::T, ScatterCfg<INSERT>::MIN_WAVES_PER_SIMD) void brick_scatter_kernel(
    const f2v *__restrict__ src, SliceGeom g, const float *__restrict__ rot, int n_rot,
    BrickGrid bg, const float *__restrict__ maxabs, f2v *__restrict__ out,
    float *__restrict__ wout, f2v *__restrict__ plane, float *__restrict__ wplane,
    u64 *__restrict__ zone_acc, const PlanePre *__restrict__ pre, int accumulate) {
  typedef Lay<BZ, BY, BX> L;
  constexpr int NV = BZ * BY * BX;
  constexpr int NS = L::SIZE;
  constexpr int kBT = ScatterCfg<INSERT>::T, kBW = ScatterCfg<INSERT>::W;
  static_assert(kItemsMax <= kBT && kRotChunk <= kBT, "one item and one rotation per thread");
  __shared__ u64 s_re[NS];
  __shared__ u64 s_im[NS];
  __shared__ u64 s_w[INSERT ? NS : 1];
  // per item (rotation, Friedel half) crossing the brick
  __shared__ int s_code[kItemsMax];     // 2 * rotation + (half == -1)
  __shared__ short s_m0[kItemsMax];     // first slice row (m = i - H/2)
  __shared__ short s_mc[kItemsMax];     // reference pixel (row m, column j)
  __shared__ short s_jc[kItemsMax];
  __shared__ int s_rowend[kItemsMax];   // inclusive prefix of row counts
  __shared__ float s_geo[kItemsMax][12]; // c(ref) - lo, s*A, s*B, 1/(s*B) (brick-local FP32)
  __shared__ RowRec s_rt[kBW][64];     // per wave: the rows of its current pass
  __shared__ int s_nitems;
  __shared__ int s_wsum[kBW];

  // The dispatcher deals consecutive blocks to the 8 XCDs in turn; in natural
  // brick order every XCD gets a similar mix of heavy (near DC) and light
  // bricks.  The extra segments of the DC-zone bricks, the heaviest units,
  // are dispatched first so that they do not form the tail.
  const int64_t b = blockIdx.x;
  if (b >= bg.n_units) return;
  const int64_t n_extra = bg.n_units - bg.n_bricks;
  const int64_t u = b < n_extra ? bg.n_bricks + b : b - n_extra;
  // unit -> (brick, rotation segment)
  int bz, by, bx, seg = 0, nseg = 1;
  bool in_zone;
  if (u < bg.n_bricks) {
    bz = (int)(u / ((int64_t)bg.nby * bg.nbx));
    const int rem = (int)(u - (int64_t)bz * bg.nby * bg.nbx);
    by = rem / bg.nbx;
    bx = rem - by * bg.nbx;
    in_zone = bz >= bg.zb0 && bz <= bg.zb1 && by >= bg.yb0 && by <= bg.yb1 && bx <= bg.xb1;
    if (in_zone) nseg = bg.nseg;
  } else {
    const int e = (int)(u - bg.n_bricks);
    const int zb = e / (bg.nseg - 1);
    seg = 1 + e - zb * (bg.nseg - 1);
    nseg = bg.nseg;
    in_zone = true;
    const int zz = zb / (bg.nzy * bg.nzx), rr = zb - zz * bg.nzy * bg.nzx;
    bz = bg.zb0 + zz;
    by = bg.yb0 + rr / bg.nzx;
    bx = rr % bg.nzx;
  }
  const int r_begin = (int)((int64_t)n_rot * seg / nseg);
  const int r_end = (int)((int64_t)n_rot * (seg + 1) / nseg);

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lo[3] = {bz * BZ, by * BY, bx * BX};
  const int hi[3] = {min(lo[0] + BZ - 1, g.vD - 1), min(lo[1] + BY - 1, g.vH - 1),
                     min(lo[2] + BX - 1, g.vW - 1)};
  const int ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};  // local max index
  const int top[3] = {g.vD - 1, g.vH - 1, g.vW - 1};
  // Footprint boxes reaching a volume face need the FP32 guard tests there.
  const bool face_lo[3] = {lo[0] <= 1, lo[1] <= 1, lo[2] <= 1};
  const bool face_hi[3] = {hi[0] >= top[0] - 1, hi[1] >= top[1] - 1, hi[2] >= top[2] - 1};
  const bool any_face = face_lo[0] || face_lo[1] || face_lo[2] || face_hi[0] || face_hi[1] || face_hi[2];
  double Lb[3], Ub[3], ctr[3], half[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Lb[d] = (double)lo[d] - 1.0 - kMargin;
    Ub[d] = (double)hi[d] + 1.0 + kMargin;
    ctr[d] = 0.5 * (Lb[d] + Ub[d]);
    half[d] = 0.5 * (Ub[d] - Lb[d]);
  }
  const double beta[3] = {g.bz, g.by, 0.0};
  const float vscale = value_scale(maxabs);
  // brick-local FP32 interval bounds, widened to cover FP32 rounding
  const float Lf[3] = {-1.0f - 1e-4f, -1.0f - 1e-4f, -1.0f - 1e-4f};
  const float Uf[3] = {(float)ext[0] + 1.0f + 1e-4f, (float)ext[1] + 1.0f + 1e-4f,
                       (float)ext[2] + 1.0f + 1e-4f};

  for (int v = tid; v < NS; v += kBT) {
    s_re[v] = 0;
    s_im[v] = 0;
    if (INSERT) s_w[v] = 0;
  }

  for (int r0 = r_begin; r0 < r_end; r0 += kRotChunk) {
    if (tid == 0) s_nitems = 0;
    __syncthreads();
    // (1) plane / box intersection, one lane per rotation
    const int r = r0 + tid;
    if (tid < kRotChunk && r < r_end) {
      const PlanePre &p = pre[r];
      const double dist = p.n[0] * (ctr[0] - beta[0]) + p.n[1] * (ctr[1] - beta[1]) +
                          p.n[2] * (ctr[2] - beta[2]);
      const double rad = fabs(p.n[0]) * half[0] + fabs(p.n[1]) * half[1] + fabs(p.n[2]) * half[2];
      if (p.nn < 0.0 || fabs(dist) <= rad + 1e-9 * p.nn) {
        double mc = 0.0, mr = 0.0, jc = 0.0, jr = 0.0;
        if (p.nn >= 0.0) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            mc += p.pm[d] * (ctr[d] - beta[d]);
            jc += p.pj[d] * (ctr[d] - beta[d]);
            mr += fabs(p.pm[d]) * half[d];
            jr += fabs(p.pj[d]) * half[d];
          }
        } else {  // degenerate (non-rotation) matrix: search every row and column
          mr = jr = 1e300;
        }
        const double mmin = (double)(-g.half_H), mmax = (double)(g.H - 1 - g.half_H);
        for (int sh = 0; sh < 2; ++sh) {
          const double s = sh ? -1.0 : 1.0;
          const double jlo = s * jc - jr, jhi = s * jc + jr;
          if (jhi < -kMargin || jlo > (double)(g.Wr - 1) + kMargin) continue;
          const double mlo = fmax(s * mc - mr, mmin), mhi = fmin(s * mc + mr, mmax);
          const int m0 = (int)ceil(mlo - kMargin), m1 = (int)floor(mhi + kMargin);
          if (m0 > m1) continue;
          // reference pixel: the one nearest the projection of the box centre
          const int mref = (int)rint(fmin(fmax(s * mc, mmin), mmax));
          const int jref = (int)rint(fmin(fmax(s * jc, 0.0), (double)(g.Wr - 1)));
          const int slot = atomicAdd(&s_nitems, 1);
          s_code[slot] = 2 * r + sh;
          s_m0[slot] = (short)m0;
          s_mc[slot] = (short)mref;
          s_jc[slot] = (short)jref;
          s_rowend[slot] = m1 - m0 + 1;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            s_geo[slot][d] =
                (float)(beta[d] + s * (p.A[d] * (double)mref + p.B[d] * (double)jref) - (double)lo[d]);
            s_geo[slot][3 + d] = (float)(s * p.A[d]);
            const float sb = (float)(s * p.B[d]);
            s_geo[slot][6 + d] = sb;
            s_geo[slot][9 + d] = fabsf(sb) > 1e-6f ? 1.0f / sb : 0.0f;
          }
        }
      }
    }
    __syncthreads();
    const int nitems = s_nitems;
    // inclusive prefix of the items' row counts (one item per thread)
    {
      int incl = tid < nitems ? s_rowend[tid] : 0;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      if (lane == 63) s_wsum[wave] = incl;
      __syncthreads();
      int base = 0;
      for (int w = 0; w < wave; ++w) base += s_wsum[w];
      incl += base;
      if (tid < nitems) s_rowend[tid] = incl;
    }
    __syncthreads();
    const int total_rows = nitems > 0 ? s_rowend[nitems - 1] : 0;
    // (2) rows of all items, 64 per wave pass
    for (int rb = wave * 64; rb < total_rows; rb += kBT) {
      const int t = rb + lane;
      int jl = 0, len = 0, code = 0, i = 0;
      float st[3] = {0.0f, 0.0f, 0.0f}, sb[3] = {0.0f, 0.0f, 0.0f};
      if (t < total_rows) {
        // item of row t: the first whose inclusive row end exceeds t.  The
        // items of one 64-row pass are few and consecutive: find the first
        // one (uniform search), then walk them (uniform loop).
        int it0 = 0;
        for (int step = kItemsMax / 2; step > 0; step >>= 1)
          if (it0 + step - 1 < nitems && s_rowend[it0 + step - 1] <= rb) it0 += step;
        it0 = __builtin_amdgcn_readfirstlane(it0);
        int it = it0;
        for (int k = it0; k < nitems; ++k) {
          const int end = s_rowend[k];
          if (end <= t) it = k + 1;
          if (end >= rb + 64) break;
        }
        const int excl = it > 0 ? s_rowend[it - 1] : 0;
        const int m = s_m0[it] + (t - excl);
        code = s_code[it];
        i = m + g.half_H;
        // brick-local FP32 coordinates relative to the item's reference pixel
        const int jc = s_jc[it];
        const float dm = (float)(m - s_mc[it]);
        float cm[3];
        float dmin = (float)(-jc), dmax = (float)(g.Wr - 1 - jc);  // j - jc range
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          sb[d] = s_geo[it][6 + d];
          cm[d] = fmaf(dm, s_geo[it][3 + d], s_geo[it][d]);
          const float inv = s_geo[it][9 + d];
          if (inv != 0.0f) {
            const float t1 = (Lf[d] - cm[d]) * inv, t2 = (Uf[d] - cm[d]) * inv;
            dmin = fmaxf(dmin, fminf(t1, t2));
            dmax = fminf(dmax, fmaxf(t1, t2));
          } else if (cm[d] < Lf[d] || cm[d] > Uf[d]) {
            dmax = -1e9f;
          }
        }
        dmin = fminf(fmaxf(dmin, -1e6f), 1e6f);  // keep the int conversions in range
        dmax = fminf(fmaxf(dmax, -1e6f), 1e6f);
        const int dl = (int)ceilf(dmin), dh = (int)floorf(dmax);
        jl = jc + dl;
        len = max(0, dh - dl + 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) st[d] = fmaf((float)dl, sb[d], cm[d]);
      }
      // (3) the rows' pixels spread over the lanes, consecutive lanes taking
      // consecutive pixels of a row (coalesced value loads)
      int incl = len;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int tt = __shfl_up(incl, off, 64);
        if (lane >= off) incl += tt;
      }
      const int total = __shfl(incl, 63, 64);
      if (len > 0) {
        RowRec rec;
        rec.jl = jl;
        rec.excl = incl - len;
        rec.i = i;
        rec.code = code;
        rec.st0 = st[0];
        rec.st1 = st[1];
        rec.st2 = st[2];
        rec.sb0 = sb[0];
        rec.sb1 = sb[1];
        rec.sb2 = sb[2];
        s_rt[wave][lane] = rec;
      }
      __builtin_amdgcn_wave_barrier();
      // Branch-free per pixel: a rejected pixel (past the row ends, outside
      // the disc, the volume or the brick) adds zero at the margin corner, so
      // the lanes of a wave never diverge on the common path.
      for (int p0 = 0; p0 < total; p0 += 64) {
        const int p = p0 + lane;
        int k = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
          const int v = __shfl(incl, k + step - 1, 64);
          if (v <= p) k += step;
        }
        const RowRec rec = s_rt[wave][k];
        const int q = p - rec.excl;
        const int j = rec.jl + q;
        const int r_i = rec.i;
        const int ri = rec.code >> 1;
        const bool neg = rec.code & 1;
        bool ok = p < total;
        if (INSERT) ok = ok && !(j == 0 && r_i < g.half_H);
        if (g.use_mask) ok = ok && in_disc(g, r_i, j);
        float lz = fmaf((float)q, rec.sb0, rec.st0);
        float ly = fmaf((float)q, rec.sb1, rec.st1);
        float lx = fmaf((float)q, rec.sb2, rec.st2);
        if (any_face) {
          const float gz = lz + (float)lo[0], gy = ly + (float)lo[1], gx = lx + (float)lo[2];
          bool exact = false, out_vol = false;
          if (face_lo[2] && gx < kGuard) (gx > -kGuard ? exact : out_vol) = true;
          if (face_lo[0] && gz < kGuard) (gz > -kGuard ? exact : out_vol) = true;
          if (face_lo[1] && gy < kGuard) (gy > -kGuard ? exact : out_vol) = true;
          if (face_hi[0] && gz > (float)top[0] - kGuard) (gz < (float)top[0] + kGuard ? exact : out_vol) = true;
          if (face_hi[1] && gy > (float)top[1] - kGuard) (gy < (float)top[1] + kGuard ? exact : out_vol) = true;
          if (face_hi[2] && gx > (float)top[2] - kGuard) (gx < (float)top[2] + kGuard ? exact : out_vol) = true;
          ok = ok && !out_vol;  // outside the volume or in the other Friedel half
          if (ok && exact) {
            const Rot rr = load_rot(rot + (int64_t)ri * 9, g);
            const Coord c = slice_coord<INSERT>(g, rr, r_i, j);
            ok = c.conj == neg && inside(g, c);
            lz = (float)(c.z - (double)lo[0]);
            ly = (float)(c.y - (double)lo[1]);
            lx = (float)(c.x - (double)lo[2]);
          }
        }
        // cell (lower corner clamped to N-2 globally, as in the gather)
        int z0 = min((int)floorf(lz), top[0] - 1 - lo[0]);
        int y0 = min((int)floorf(ly), top[1] - 1 - lo[1]);
        int x0 = min((int)floorf(lx), top[2] - 1 - lo[2]);
        ok = ok && z0 >= -1 && z0 <= ext[0] && y0 >= -1 && y0 <= ext[1] && x0 >= -1 && x0 <= ext[2];
        const float tz = lz - (float)z0, ty = ly - (float)y0, tx = lx - (float)x0;
        if (!ok) z0 = y0 = x0 = -1;
        const int64_t src_off = ok ? (int64_t)ri * g.P + (int64_t)r_i * g.Wr + j : 0;
        f2v val = src[src_off];
        val.y = neg ? -val.y : val.y;
        const float vs = ok ? vscale : 0.0f;
        val *= vs;
        const float wz[2] = {1.0f - tz, tz}, wy[2] = {1.0f - ty, ty}, wx[2] = {1.0f - tx, tx};
        const float wsc = ok ? kWeightScale : 0.0f;
        const int e = L::at(z0, y0, x0);
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
          for (int dy = 0; dy < 2; ++dy) {
            const float wzy = wz[dz] * wy[dy];
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
              const float w = wzy * wx[dx];
              const int o = e + dz * L::PZ + dy * L::PY + dx;
              lds_add(&s_re[o], w * val.x);
              lds_add(&s_im[o], w * val.y);
              if (INSERT) lds_add(&s_w[o], w * wsc);
            }
          }
      }
      __builtin_amdgcn_wave_barrier();  // the next pass rewrites this wave's row table
    }
    __syncthreads();
  }

  // (4) flush
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  const float inv_v = 1.0f / vscale;
  if (in_zone) {
    const int zb = ((bz - bg.zb0) * bg.nzy + (by - bg.yb0)) * bg.nzx + bx;
    u64 *acc = zone_acc + (int64_t)zb * (INSERT ? 3 : 2) * NV;
    for (int v = tid; v < NV; v += kBT) {
      const int ls = L::at(v / (BY * BX), (v / BX) % BY, v % BX);
      if (s_re[ls]) atomicAdd(acc + v, s_re[ls]);
      if (s_im[ls]) atomicAdd(acc + NV + v, s_im[ls]);
      if (INSERT && s_w[ls]) atomicAdd(acc + 2 * NV + v, s_w[ls]);
    }
    return;
  }
  for (int v = tid; v < NV; v += kBT) {
    const int lz = v / (BY * BX), ly = (v / BX) % BY, lx = v % BX;
    const int z = lo[0] + lz, y = lo[1] + ly, x = lo[2] + lx;
    if (z > hi[0] || y > hi[1] || x > hi[2]) continue;
    const int ls = L::at(lz, ly, lx);
    f2v val = {(float)(long long)s_re[ls] * inv_v, (float)(long long)s_im[ls] * inv_v};
    const float w = INSERT ? (float)(long long)s_w[ls] * (1.0f / kWeightScale) : 0.0f;
    if (INSERT && x == 0) {
      plane[(int64_t)z * g.vH + y] = val;
      wplane[(int64_t)z * g.vH + y] = w;
      continue;
    }
    const int64_t e = (int64_t)z * sz + (int64_t)y * sy + x;
    if (accumulate) val += out[e];
    out[e] = val;
    if (INSERT) wout[e] = w + (accumulate ? wout[e] : 0.0f);
  }
}

// Converts the zone bricks' 64-bit partial sums and writes them like a flush.
template <int BZ, int BY, int BX, bool INSERT>
__global__ __launch_bounds__(kT) void zone_combine_kernel(
    SliceGeom g, BrickGrid bg, const float *__restrict__ maxabs, const u64 *__restrict__ zone_acc,
    f2v *__restrict__ out, float *__restrict__ wout, f2v *__restrict__ plane,
    float *__restrict__ wplane, int accumulate) {
  constexpr int NV = BZ * BY * BX;
  const int zb = blockIdx.x;
  if (zb >= bg.n_zone) return;
  const int zz = zb / (bg.nzy * bg.nzx), rr = zb - zz * bg.nzy * bg.nzx;
  const int lo[3] = {(bg.zb0 + zz) * BZ, (bg.yb0 + rr / bg.nzx) * BY, (rr % bg.nzx) * BX};
  const float inv_v = 1.0f / value_scale(maxabs);
  const u64 *acc = zone_acc + (int64_t)zb * (INSERT ? 3 : 2) * NV;
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  for (int v = threadIdx.x; v < NV; v += kT) {
    const int z = lo[0] + v / (BY * BX), y = lo[1] + (v / BX) % BY, x = lo[2] + v % BX;
    if (z >= g.vD || y >= g.vH || x >= g.vW) continue;
    f2v val = {(float)(long long)acc[v] * inv_v, (float)(long long)acc[NV + v] * inv_v};
    const float w = INSERT ? (float)(long long)acc[2 * NV + v] * (1.0f / kWeightScale) : 0.0f;
    if (INSERT && x == 0) {
      plane[(int64_t)z * g.vH + y] = val;
      wplane[(int64_t)z * g.vH + y] = w;
      continue;
    }
    const int64_t e = (int64_t)z * sz + (int64_t)y * sy + x;
    if (accumulate) val += out[e];
    out[e] = val;
    if (INSERT) wout[e] = w + (accumulate ? wout[e] : 0.0f);
  }
}

// x=0-plane Friedel mirror of the insertion, from the plane of direct splats.
__global__ __launch_bounds__(kT) void mirror_plane_kernel(const f2v *__restrict__ plane,
                                                          const float *__restrict__ wplane, int D,
                                                          int H, int Wr, f2v *__restrict__ out,
                                                          float *__restrict__ wout, int accumulate) {
  const int idx = blockIdx.x * kT + threadIdx.x;
  if (idx >= D * H) return;
  const int z = idx / H, y = idx - z * H;
  const int mz = (D - z) % D, my = (H - y) % H;
  f2v val = plane[idx];
  float w = wplane[idx];
  if (!(mz == D / 2 && my == H / 2)) {
    const f2v m = plane[mz * H + my];
    val.x += m.x;
    val.y -= m.y;
    w += wplane[mz * H + my];
  }
  const int64_t e = ((int64_t)z * H + y) * Wr;
  if (accumulate) {
    val += out[e];
    w += wout[e];
  }
  out[e] = val;
  wout[e] = w;
}

// max(|re|, |im|) over n complex values into *res (float bits, res >= 0).
// Reads 16 bytes (two complex values) per lane and load, four loads in
// flight per lane; an odd tail element is read by thread 0.
__global__ __launch_bounds__(kT) void maxabs_kernel(const f2v *__restrict__ x, int64_t n,
                                                    unsigned int *__restrict__ res) {
  const float4 *x4 = reinterpret_cast<const float4 *>(x);
  const int64_t n4 = n >> 1;
  const int64_t stride = (int64_t)gridDim.x * kT;
  float m = 0.0f;
  int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x;
  for (; p + 3 * stride < n4; p += 4 * stride) {
    const float4 a = x4[p], b = x4[p + stride], c = x4[p + 2 * stride], d = x4[p + 3 * stride];
    m = fmaxf(m, fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))),
                       fmaxf(fmaxf(fabsf(b.x), fabsf(b.y)), fmaxf(fabsf(b.z), fabsf(b.w)))));
    m = fmaxf(m, fmaxf(fmaxf(fmaxf(fabsf(c.x), fabsf(c.y)), fmaxf(fabsf(c.z), fabsf(c.w))),
                       fmaxf(fmaxf(fabsf(d.x), fabsf(d.y)), fmaxf(fabsf(d.z), fabsf(d.w)))));
  }
  for (; p < n4; p += stride) {
    const float4 a = x4[p];
    m = fmaxf(m, fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))));
  }
  if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
    const f2v v = x[n - 1];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(res, __float_as_uint(m));
}

// Scalar fallback for a source that is not 16-byte aligned.
__global__ __launch_bounds__(kT) void maxabs_scalar_kernel(const f2v *__restrict__ x, int64_t n,
                                                           unsigned int *__restrict__ res) {
  float m = 0.0f;
  for (int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x; p < n; p += (int64_t)gridDim.x * kT) {
    const f2v v = x[p];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(res, __float_as_uint(m));
}

template <typename T>
__global__ __launch_bounds__(kT) void zero_words_kernel(T *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kT)
    p[i] = (T)0;
}

template <int BZ, int BY, int BX>
BrickGrid make_grid(const SliceGeom &g, int nseg, int zone_r, int zone_rx) {
  BrickGrid bg;
  bg.nbz = (g.vD + BZ - 1) / BZ;
  bg.nby = (g.vH + BY - 1) / BY;
  bg.nbx = (g.vW + BX - 1) / BX;
  bg.n_bricks = bg.nbz * bg.nby * bg.nbx;
  bg.nseg = std::max(nseg, 1);
  // zone of bricks around the DC coordinate (bz, by, 0), which all slices cross
  const int cz = std::min(std::max((int)std::floor(g.bz / BZ), 0), bg.nbz - 1);
  const int cy = std::min(std::max((int)std::floor(g.by / BY), 0), bg.nby - 1);
  bg.zb0 = std::max(cz - zone_r - 1, 0);
  bg.zb1 = std::min(cz + zone_r, bg.nbz - 1);
  bg.yb0 = std::max(cy - zone_r - 1, 0);
  bg.yb1 = std::min(cy + zone_r, bg.nby - 1);
  bg.xb1 = std::min(zone_rx, bg.nbx - 1);
  if (bg.nseg == 1) {  // no split: an empty zone
    bg.zb0 = bg.yb0 = 1;
    bg.zb1 = bg.yb1 = 0;
    bg.xb1 = -1;
  }
  bg.nzy = std::max(bg.yb1 - bg.yb0 + 1, 0);
  bg.nzx = bg.xb1 + 1;
  bg.n_zone = std::max(bg.zb1 - bg.zb0 + 1, 0) * bg.nzy * bg.nzx;
  bg.n_units = bg.n_bricks + bg.n_zone * (bg.nseg - 1);
  bg.n_units_padded = (bg.n_units + 7) & ~7;
  return bg;
}

template <int BZ, int BY, int BX, bool INSERT>
int launch_brick(const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                 float *wout, int accumulate, hipStream_t s) {
  constexpr int NV = BZ * BY * BX;
  // split the DC-zone bricks when there are enough rotations to share
  const int nseg = n_rot >= 512 ? 8 : (n_rot >= 128 ? 4 : 1);
  const BrickGrid bg = make_grid<BZ, BY, BX>(g, nseg, 2, 2);
  const size_t plane_elems = (size_t)g.vD * g.vH;
  const size_t zone_words = (size_t)bg.n_zone * (INSERT ? 3 : 2) * NV;
  const size_t off_plane = 256;
  const size_t off_wplane = off_plane + (INSERT ? plane_elems * sizeof(f2v) : 0);
  const size_t off_zone =
      (off_wplane + (INSERT ? plane_elems * sizeof(float) : 0) + 255) & ~(size_t)255;
  const size_t off_pre = (off_zone + zone_words * sizeof(u64) + 255) & ~(size_t)255;
  const size_t bytes = off_pre + (size_t)n_rot * sizeof(PlanePre);
  void *scratch = nullptr;
  hipError_t e = hipMallocAsync(&scratch, bytes, s);
  if (e != hipSuccess) {
    set_error(std::string("hipMallocAsync: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  char *base = (char *)scratch;
  unsigned int *maxabs = (unsigned int *)base;
  f2v *plane = INSERT ? (f2v *)(base + off_plane) : nullptr;
  float *wplane = INSERT ? (float *)(base + off_wplane) : nullptr;
  u64 *zone = zone_words ? (u64 *)(base + off_zone) : nullptr;
  PlanePre *pre = (PlanePre *)(base + off_pre);
  const int64_t n = n_rot * (int64_t)g.P;
  hipLaunchKernelGGL(zero_words_kernel<unsigned int>, dim3(1), dim3(kT), 0, s, maxabs, (int64_t)1);
  if (((uintptr_t)src & 15) == 0)
    hipLaunchKernelGGL(maxabs_kernel, dim3(2048), dim3(kT), 0, s, src, n, maxabs);
  else
    hipLaunchKernelGGL(maxabs_scalar_kernel, dim3(2048), dim3(kT), 0, s, src, n, maxabs);
  if (zone_words)
    hipLaunchKernelGGL(zero_words_kernel<u64>,
                       dim3((int)std::min<size_t>((zone_words + kT - 1) / kT, 4096)), dim3(kT), 0,
                       s, zone, (int64_t)zone_words);
  hipLaunchKernelGGL(plane_prep_kernel, dim3((unsigned)((n_rot + kT - 1) / kT)), dim3(kT), 0, s, g, rot,
                     (int)n_rot, pre);
  int st = check_launch("maxabs / zero / plane_prep");
  if (!st) {
    hipLaunchKernelGGL((brick_scatter_kernel<BZ, BY, BX, INSERT>), dim3(bg.n_units_padded),
                       dim3(ScatterCfg<INSERT>::T), 0, s, src, g, rot, (int)n_rot, bg, (const float *)maxabs, out,
                       wout, plane, wplane, zone, pre, accumulate);
    st = check_launch("brick_scatter_kernel");
  }
  if (!st && bg.n_zone) {
    hipLaunchKernelGGL((zone_combine_kernel<BZ, BY, BX, INSERT>), dim3(bg.n_zone), dim3(kT), 0, s,
                       g, bg, (const float *)maxabs, (const u64 *)zone, out, wout, plane, wplane,
                       accumulate);
    st = check_launch("zone_combine_kernel");
  }
  if (!st && INSERT) {
    const int np = (int)plane_elems;
    hipLaunchKernelGGL(mirror_plane_kernel, dim3((np + kT - 1) / kT), dim3(kT), 0, s, plane,
                       wplane, g.vD, g.vH, g.vW, out, wout, accumulate);
    st = check_launch("mirror_plane_kernel");
  }
  e = hipFreeAsync(scratch, s);
  if (!st && e != hipSuccess) {
    set_error(std::string("hipFreeAsync: ") + hipGetErrorString(e));
    st = TT_ERR_HIP;
  }
  return st;
}

}  // namespace

// The fixed-point brick scatter is bitwise reproducible, so the deterministic
// flag needs no separate path.
int brick_scatter_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot,
                          f2v *gV, int /*deterministic*/, int accumulate, hipStream_t s) {
  return launch_brick<16, 16, 8, false>(g, gS, rot, n_rot, gV, nullptr, accumulate, s);
}

int brick_scatter_insert(const SliceGeom &g, const f2v *slices, const float *rot, int64_t n_rot,
                         f2v *vol, float *w, int /*deterministic*/, int accumulate, hipStream_t s) {
  return launch_brick<16, 16, 8, true>(g, slices, rot, n_rot, vol, w, accumulate, s);
}

}  // namespace tt
