// Brick engine for gfx950: the central-slice gather (extraction) and scatter
// (extraction adjoint, insertion) organised around bricks of the volume.
//
// A workgroup owns one brick of BZ x BY x BX trilinear cells (a "unit" is a
// brick and a segment of the rotations).  A slice pixel belongs to exactly one
// brick: the one holding its canonical cell (lower corner, clamped to N-2; see
// tt_common.h), so every pixel inside the volume is processed exactly once.
//  gather : the brick's voxels (cells + one upper halo voxel per axis) are
//           copied into LDS once and every slice crossing the brick is served
//           from there; pixels outside the volume or the fftfreq_max disc are
//           written by a row-interval zero pass that runs first.
//  scatter: the brick accumulates in LDS with ds_add_f32 (8 corners per
//           pixel, the upper corners land in the halo), then adds its voxels
//           to the output with global float atomics.
//
// Inside a workgroup the waves are independent: each takes a contiguous share
// of the unit's rotations and runs
//  (1) discovery: one lane per rotation tests the slice plane against the
//      brick's coordinate box and records, for each Friedel half that reaches
//      it, the range of slice rows (item);
//  (2) rows: one lane per row evaluates the canonical row constants and the
//      interval of columns whose coordinates can lie in the box (FP64 slab
//      constraints, widened by 1e-6 voxel);
//  (3) pixels: the rows' columns are cut into chunks of CH pixels; a wave
//      prefix sum hands one chunk to each lane, consecutive lanes taking
//      consecutive chunks of a row.  Each pixel evaluates its canonical FP64
//      coordinate and keeps it only if its cell is in the brick.
// No workgroup barrier is taken between the tile load (or clear) and the
// flush, so the 8 waves of a block overlap their latencies freely.
//
// Units are ordered heaviest first (estimated from the fraction of uniformly
// oriented planes that cross a brick at its distance from DC); heavy bricks
// near DC are split into rotation segments so no unit dominates the tail.
//
// Reference behaviour (packages/primitives/torch-fourier-slice/src/torch_fourier_slice):
//   extraction  slice_extraction/_extract_central_slices_rfft_3d.py:84-222
//   insertion   slice_insertion/_insert_central_slices_rfft_3d.py:9-106,213-406
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_engine.h"

namespace tt {

namespace {

enum Mode { kGather = 0, kAdjoint = 1, kInsert = 2 };

constexpr double kBoxMargin = 1e-6;  // widening of the candidate box (voxels)
constexpr int kDiscRot = 32;         // rotations tested per discovery pass (<= 64 items)
constexpr int kItems = 2 * kDiscRot;
constexpr int kRows = 32;            // rows per row pass

// Per-rotation plane data, shared by all bricks.
struct PlaneRec {
  double U[3], V[3];    // canonical constants (tt_common.h)
  double iV[3];         // 1 / V_d, 0 where |V_d| <= 1e-30 (constant along a row)
  double n[3];          // normal of span(A, B), A = U / H (per row), B = V (per column)
  double pm[3], pj[3];  // rows of the pseudo-inverse: (m, j) of a point of the plane
  double nn;            // |n|; negative for a degenerate matrix (no pseudo-inverse)
  int tie, pad;
};

__global__ __launch_bounds__(256) void plane_rec_kernel(SliceGeom g, const float *__restrict__ rot,
                                                        int n_rot, PlaneRec *__restrict__ out) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n_rot) return;
  const Rot rr = load_rot(rot + (int64_t)r * 9, g);
  PlaneRec p;
  double A[3], B[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    p.U[d] = rr.U[d];
    p.V[d] = rr.V[d];
    p.iV[d] = fabs(rr.V[d]) > 1e-30 ? 1.0 / rr.V[d] : 0.0;
    A[d] = rr.U[d] * g.inv_H;
    B[d] = rr.V[d];
  }
  p.n[0] = A[1] * B[2] - A[2] * B[1];
  p.n[1] = A[2] * B[0] - A[0] * B[2];
  p.n[2] = A[0] * B[1] - A[1] * B[0];
  p.nn = sqrt(p.n[0] * p.n[0] + p.n[1] * p.n[1] + p.n[2] * p.n[2]);
  const double AA = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  const double BB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  const double AB = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
  const double det = AA * BB - AB * AB;
  if (det > 1e-20 * AA * BB && AA > 0.0 && BB > 0.0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      p.pm[d] = (BB * A[d] - AB * B[d]) / det;
      p.pj[d] = (AA * B[d] - AB * A[d]) / det;
    }
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) p.pm[d] = p.pj[d] = 0.0;
    p.nn = -1.0;
  }
  p.tie = rr.tie ? 1 : 0;
  p.pad = 0;
  out[r] = p;
}

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

__device__ __forceinline__ int rank_below(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// First lane k in [0, NL) whose inclusive prefix `incl` exceeds p (NL a power of 2).
template <int NL>
__device__ __forceinline__ int find_segment(int incl, int p) {
  int k = 0;
#pragma unroll
  for (int step = NL / 2; step > 0; step >>= 1) {
    const int v = __shfl(incl, k + step - 1, 64);
    if (v <= p) k += step;
  }
  return k;
}

// ---------------------------------------------------------------------------
// brick units
// ---------------------------------------------------------------------------
struct BrickGrid {
  int nbz, nby, nbx;  // bricks of cells (cells run 0 .. N-2 per axis)
};

struct Unit {
  int brick;
  int seg;   // segment index << 16 | number of segments (0: an empty brick)
  int slot;  // first partial slot of a split scatter brick, else -1
};

struct SplitRec {
  int brick, nseg, slot;  // a scatter brick split into nseg rotation segments
};

template <int BZ, int BY, int BX, int MODE>
struct Cfg {
  static constexpr int NW = 8;                          // waves per block
  static constexpr int T = NW * 64;
  static constexpr int CH = (MODE == kGather) ? 4 : 2;  // pixels per chunk
  // gather: cells lo..hc and their corners (B + 1 voxels per axis);
  // scatter: owned voxels lo..hi with a one-voxel margin (B + 2 per axis)
  static constexpr int OFF = (MODE == kGather) ? 0 : 1;
  static constexpr int SX = BX + 1 + OFF, SY = BY + 1 + OFF, SZ = BZ + 1 + OFF;
  static constexpr int NT = SX * SY * SZ;
  static constexpr int NACC = (MODE == kGather) ? 0 : ((MODE == kInsert) ? 3 : 2);
  static constexpr int NV = BZ * BY * BX;  // owned voxels of a scatter brick
};

struct ItemRec {
  int r;     // rotation (relative to the launch)
  int m0;    // first row, as m = i - H//2
  int nrow;  // number of rows
  int neg;   // Friedel half: 0 -> s = +1, 1 -> s = -1
};

struct RowRec {
  double C[3];     // canonical row constants fma(s U_d, fy, b_d) (x: s U_x fy)
  double sV[3];    // s V_d
  long long base;  // r * P + i * Wr
  int jl;          // first candidate column
  int info;        // i | neg << 16 | tie << 17
};

// Fixed-point scale of the scatter: values are multiplied by
// 2^(30 - headroom - e), with 2^e > the sampled max|v|, so a value up to
// 2^headroom times the sampled max still converts to a 32-bit integer; a
// larger (or non-finite) value raises *flag and the call is redone with the
// exact maximum (headroom 0).
struct ScaleArgs {
  const unsigned *maxbits;  // float bits of max|v| (sampled or exact)
  unsigned *flag;           // set when a value exceeds the limit
  int headroom;             // bits of headroom above the sampled maximum
  int guard;                // 1: run only if *flag is set (the exact redo)
};

__device__ __forceinline__ void scale_of(const ScaleArgs &sa, float &scale, float &limit) {
  const float m = __uint_as_float(*sa.maxbits);
  int e = 0;
  if (m > 0.0f && m < 3.0e38f) frexpf(m, &e);
  scale = ldexpf(1.0f, 30 - sa.headroom - e);
  limit = ldexpf(1.0f, e + sa.headroom);
}

constexpr float kWScale = 1073741824.0f;  // 2^30: fixed-point scale of the weights

typedef unsigned long long u64;

__device__ __forceinline__ void lds_add_fixed(u64 *p, float x) {
  const int q = __float2int_rn(x);
  atomicAdd(p, (u64)(long long)q);
}

// ---------------------------------------------------------------------------
// The brick kernel (gather, adjoint or insertion)
// ---------------------------------------------------------------------------
template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__((Cfg<BZ, BY, BX, MODE>::T)) void brick_kernel(
    const f2v *__restrict__ src, SliceGeom g, BrickGrid bg, const PlaneRec *__restrict__ pre,
    const Unit *__restrict__ units, int n_rot, f2v *__restrict__ out, float *__restrict__ wout,
    f2v *__restrict__ plane, float *__restrict__ wplane, u64 *__restrict__ segbuf, ScaleArgs sa,
    int accumulate) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int NW = C::NW, CH = C::CH, SX = C::SX, SY = C::SY, NT = C::NT, OFF = C::OFF;
  constexpr bool SCATTER = MODE != kGather;
  // gather: NT complex voxels; scatter: NACC fixed-point planes of NT words
  constexpr int NWORDS = SCATTER ? C::NACC * NT : NT;
  __shared__ u64 s_tile[NWORDS];
  __shared__ ItemRec s_item[NW][kItems];
  __shared__ RowRec s_row[NW][kRows];

  if (SCATTER && sa.guard && *sa.flag == 0u) return;  // exact redo not needed
  const Unit un = units[blockIdx.x];
  const int brick = un.brick;
  const int seg = un.seg >> 16, nseg_raw = un.seg & 0xffff;
  const int nseg = max(nseg_raw, 1);  // an empty brick (no slice reaches it) only writes zeros
  const int bz = brick / (bg.nby * bg.nbx);
  const int brem = brick - bz * bg.nby * bg.nbx;
  const int by = brem / bg.nbx, bx = brem - by * bg.nbx;
  const int lo0 = bz * BZ, lo1 = by * BY, lo2 = bx * BX;
  // candidate cells [cl, ch]: gather, the brick's own cells; scatter, every
  // cell with a corner among the owned voxels lo..hi
  int cl0, cl1, cl2, ch0, ch1, ch2;
  if (SCATTER) {
    cl0 = max(lo0 - 1, 0);
    cl1 = max(lo1 - 1, 0);
    cl2 = max(lo2 - 1, 0);
    ch0 = min(lo0 + BZ - 1, g.vD - 2);
    ch1 = min(lo1 + BY - 1, g.vH - 2);
    ch2 = min(lo2 + BX - 1, g.vW - 2);
  } else {
    cl0 = lo0;
    cl1 = lo1;
    cl2 = lo2;
    ch0 = min(lo0 + BZ - 1, g.vD - 2);
    ch1 = min(lo1 + BY - 1, g.vH - 2);
    ch2 = min(lo2 + BX - 1, g.vW - 2);
  }
  const unsigned ext0 = (unsigned)(ch0 - cl0), ext1 = (unsigned)(ch1 - cl1), ext2 = (unsigned)(ch2 - cl2);
  const int to0 = lo0 - OFF, to1 = lo1 - OFF, to2 = lo2 - OFF;  // tile origin (voxels)
  const bool top0 = ch0 == g.vD - 2, top1 = ch1 == g.vH - 2, top2 = ch2 == g.vW - 2;
  const bool conj_check = cl2 == 0;  // only bricks at x = 0 see both Friedel halves
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  float vscale = 1.0f, vlimit = 3.0e38f;
  if (SCATTER) scale_of(sa, vscale, vlimit);
  bool bad = false;  // a value beyond the fixed-point limit was seen

  if (!SCATTER) {
    f2v *tv = reinterpret_cast<f2v *>(s_tile);
    for (int v = tid; v < NT; v += C::T) {
      const int z = v / (SY * SX), y = (v / SX) % SY, x = v % SX;
      const int gz = lo0 + z, gy = lo1 + y, gx = lo2 + x;
      f2v val = {0.0f, 0.0f};
      if (gz < g.vD && gy < g.vH && gx < g.vW) val = src[((int64_t)gz * g.vH + gy) * g.vW + gx];
      tv[v] = val;
    }
  } else {
    for (int v = tid; v < NWORDS; v += C::T) s_tile[v] = 0ull;
  }
  __syncthreads();

  const double Lb[3] = {(double)cl0 - kBoxMargin, (double)cl1 - kBoxMargin, (double)cl2 - kBoxMargin};
  const double Ub[3] = {(double)(ch0 + 1) + kBoxMargin, (double)(ch1 + 1) + kBoxMargin,
                        (double)(ch2 + 1) + kBoxMargin};
  const double beta[3] = {g.bz, g.by, 0.0};
  double ctr[3], half[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ctr[d] = 0.5 * (Lb[d] + Ub[d]);
    half[d] = 0.5 * (Ub[d] - Lb[d]);
  }
  const double mmin = (double)(-g.half_H), mmax = (double)(g.H - 1 - g.half_H);

  // this wave's share of the unit's rotations
  const int64_t ub = (int64_t)n_rot * seg / nseg, ue = (int64_t)n_rot * (seg + 1) / nseg;
  const int r_beg = (int)(ub + (ue - ub) * wave / NW);
  const int r_end = (int)(ub + (ue - ub) * (wave + 1) / NW);

  ItemRec *items = s_item[wave];
  RowRec *rows = s_row[wave];
  const bool have_cells = nseg_raw > 0 && ch0 >= cl0 && ch1 >= cl1 && ch2 >= cl2;

  for (int rd = r_beg; have_cells && rd < r_end; rd += kDiscRot) {
    // (1) discovery: lanes 0..kDiscRot-1 test one rotation each
    bool h0 = false, h1 = false;
    int m0a = 0, m1a = -1, m0b = 0, m1b = -1;
    const int r = rd + lane;
    if (lane < kDiscRot && r < r_end) {
      const PlaneRec &p = pre[r];
      const double dist = p.n[0] * (ctr[0] - beta[0]) + p.n[1] * (ctr[1] - beta[1]) +
                          p.n[2] * (ctr[2] - beta[2]);
      const double rad = fabs(p.n[0]) * half[0] + fabs(p.n[1]) * half[1] + fabs(p.n[2]) * half[2];
      if (p.nn < 0.0 || fabs(dist) <= rad + 1e-9 * p.nn) {
        double mc = 0.0, mr = 1e300, jc = 0.0, jr = 1e300;
        if (p.nn >= 0.0) {
          mr = jr = 0.0;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            mc += p.pm[d] * (ctr[d] - beta[d]);
            jc += p.pj[d] * (ctr[d] - beta[d]);
            mr += fabs(p.pm[d]) * half[d];
            jr += fabs(p.pj[d]) * half[d];
          }
          mr += 1e-6 * (1.0 + fabs(mc));
          jr += 1e-6 * (1.0 + fabs(jc));
        }
#pragma unroll
        for (int sh = 0; sh < 2; ++sh) {
          const double s = sh ? -1.0 : 1.0;
          if (s * jc + jr < 0.0 || s * jc - jr > (double)(g.Wr - 1)) continue;
          const double mlo = fmax(s * mc - mr, mmin), mhi = fmin(s * mc + mr, mmax);
          if (mlo > mhi) continue;
          const int m0 = (int)ceil(mlo), m1 = (int)floor(mhi);
          if (m0 > m1) continue;
          if (sh == 0) {
            h0 = true;
            m0a = m0;
            m1a = m1;
          } else {
            h1 = true;
            m0b = m0;
            m1b = m1;
          }
        }
      }
    }
    const unsigned long long b0 = __ballot(h0), b1 = __ballot(h1);
    const int n0 = __popcll(b0);
    const int nitems = n0 + __popcll(b1);
    if (h0) items[rank_below(b0)] = ItemRec{r, m0a, m1a - m0a + 1, 0};
    if (h1) items[n0 + rank_below(b1)] = ItemRec{r, m0b, m1b - m0b + 1, 1};
    __builtin_amdgcn_wave_barrier();
    if (nitems == 0) continue;

    // (2) rows of all items: lane = item for the prefix
    const int nrow_l = lane < nitems ? items[lane].nrow : 0;
    const int rincl = wave_incl_scan(nrow_l);
    const int total_rows = __shfl(rincl, 63, 64);
    for (int rb = 0; rb < total_rows; rb += kRows) {
      int len = 0;
      const int t = rb + lane;
      // item of row t: searched with every lane active (cross-lane reads)
      const int k = find_segment<64>(rincl, min(t, total_rows - 1));
      const int excl_raw = __shfl(rincl, k > 0 ? k - 1 : 0, 64);
      if (lane < kRows && t < total_rows) {
        const ItemRec it = items[k];
        const int excl = k > 0 ? excl_raw : 0;
        const int m = it.m0 + (t - excl);
        const int i = m + g.half_H;
        const double s = it.neg ? -1.0 : 1.0;
        const PlaneRec &p = pre[it.r];
        const double fy = canon_fy(g, i);
        RowRec rec;
        rec.C[0] = __fma_rn(s * p.U[0], fy, g.bz);
        rec.C[1] = __fma_rn(s * p.U[1], fy, g.by);
        rec.C[2] = s * __dmul_rn(p.U[2], fy);
        double jlo = 0.0, jhi = (double)(g.Wr - 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          rec.sV[d] = s * p.V[d];
          if (p.iV[d] != 0.0) {
            const double inv = s * p.iV[d];
            const double t1 = (Lb[d] - rec.C[d]) * inv, t2 = (Ub[d] - rec.C[d]) * inv;
            jlo = fmax(jlo, fmin(t1, t2));
            jhi = fmin(jhi, fmax(t1, t2));
          } else if (rec.C[d] < Lb[d] || rec.C[d] > Ub[d]) {
            jhi = -1.0;
          }
        }
        int jl = 0, jh = -1;
        if (jlo <= jhi) {
          jl = (int)ceil(jlo - 1e-6);
          jh = (int)floor(jhi + 1e-6);
          jl = max(jl, 0);
          jh = min(jh, g.Wr - 1);
        }
        if (MODE == kInsert && i < g.half_H) jl = max(jl, 1);  // x = 0, y < 0 is skipped
        len = max(0, jh - jl + 1);
        rec.base = (long long)it.r * g.P + (long long)i * g.Wr;
        rec.jl = jl;
        rec.info = i | (it.neg << 16) | (p.tie << 17);
        rows[lane] = rec;
      }
      const int nch = (len + CH - 1) / CH;
      const int cincl = wave_incl_scan(nch);
      const int total_ch = __shfl(cincl, 63, 64);
      __builtin_amdgcn_wave_barrier();

      // (3) chunks of CH pixels, one per lane
      for (int pc = 0; pc < total_ch; pc += 64) {
        const int pidx = pc + lane;
        const int k = find_segment<kRows>(cincl, pidx);
        const int cexcl_raw = __shfl(cincl, k > 0 ? k - 1 : 0, 64);
        const int cexcl = k > 0 ? cexcl_raw : 0;
        const int len_k = __shfl(len, k, 64);
        const bool live = pidx < total_ch;
        const int q0 = (pidx - cexcl) * CH;
        const RowRec rec = rows[k];
        const int i = rec.info & 0xffff;
        const bool neg = (rec.info >> 16) & 1;
        const bool tie = (rec.info >> 17) & 1;
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const int jj = q0 + q;
          const int j = rec.jl + jj;
          bool own = live && jj < len_k;
          const double jd = (double)j;
          double cz = __fma_rn(jd, rec.sV[0], rec.C[0]);
          double cy = __fma_rn(jd, rec.sV[1], rec.C[1]);
          double cx = __fma_rn(jd, rec.sV[2], rec.C[2]);
          if (conj_check) {
            const double cx1 = neg ? -cx : cx;
            own = own && (canon_conj<MODE == kInsert>(g, tie, cx1) == neg);
          }
          double fz = floor(cz), fy_ = floor(cy), fx = floor(cx);
          double tz = cz - fz, ty = cy - fy_, tx = cx - fx;
          const bool near = (tz < 2e-9) | (tz > 1.0 - 2e-9) | (ty < 2e-9) | (ty > 1.0 - 2e-9) |
                            (tx < 2e-9) | (tx > 1.0 - 2e-9);
          if (near) {  // rare: the canonical snap can move the coordinate
            cz = snap(cz);
            cy = snap(cy);
            cx = snap(cx);
            fz = floor(cz);
            fy_ = floor(cy);
            fx = floor(cx);
            tz = cz - fz;
            ty = cy - fy_;
            tx = cx - fx;
          }
          int z0 = (int)fz, y0 = (int)fy_, x0 = (int)fx;
          // c = N-1 exactly: the cell is clamped to N-2 with fraction 1
          if (top0 && z0 == g.vD - 1 && tz == 0.0) { z0 = g.vD - 2; tz = 1.0; }
          if (top1 && y0 == g.vH - 1 && ty == 0.0) { y0 = g.vH - 2; ty = 1.0; }
          if (top2 && x0 == g.vW - 1 && tx == 0.0) { x0 = g.vW - 2; tx = 1.0; }
          own = own && (unsigned)(z0 - cl0) <= ext0 && (unsigned)(y0 - cl1) <= ext1 &&
                (unsigned)(x0 - cl2) <= ext2;
          if (g.use_mask) own = own && in_disc(g, i, j);
          const int e = own ? ((z0 - to0) * SY + (y0 - to1)) * SX + (x0 - to2) : 0;
          const float ftz = (float)tz, fty = (float)ty, ftx = (float)tx;
          if (MODE == kGather) {
            const f2v *b = reinterpret_cast<const f2v *>(s_tile) + e;
            f2v val = trilerp(b[0], b[1], b[SX], b[SX + 1], b[SY * SX], b[SY * SX + 1],
                              b[SY * SX + SX], b[SY * SX + SX + 1], ftz, fty, ftx);
            if (neg) val.y = -val.y;
            if (own) out[rec.base + j] = val;
          } else if (own) {
            f2v v = src[rec.base + j];
            if (neg) v.y = -v.y;
            bad = bad || !(fabsf(v.x) < vlimit && fabsf(v.y) < vlimit);
            v.x *= vscale;
            v.y *= vscale;
            const float wz[2] = {1.0f - ftz, ftz}, wy[2] = {1.0f - fty, fty}, wx[2] = {1.0f - ftx, ftx};
#pragma unroll
            for (int dz = 0; dz < 2; ++dz)
#pragma unroll
              for (int dy = 0; dy < 2; ++dy) {
                const float wzy = wz[dz] * wy[dy];
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                  const float w = wzy * wx[dx];
                  const int o = e + dz * SY * SX + dy * SX + dx;
                  lds_add_fixed(&s_tile[o], w * v.x);
                  lds_add_fixed(&s_tile[NT + o], w * v.y);
                  if (MODE == kInsert) lds_add_fixed(&s_tile[2 * NT + o], w * kWScale);
                }
              }
          }
        }
      }
      __builtin_amdgcn_wave_barrier();  // the next row pass rewrites this wave's row table
    }
    __builtin_amdgcn_wave_barrier();  // the next discovery pass rewrites the item table
  }

  if (!SCATTER) return;
  if (__ballot(bad) && lane == 0) atomicOr(sa.flag, 1u);
  __syncthreads();
  // flush the owned voxels lo..hi (tile coordinates 1..B per axis)
  const int hi0 = min(lo0 + BZ - 1, g.vD - 1), hi1 = min(lo1 + BY - 1, g.vH - 1),
            hi2 = min(lo2 + BX - 1, g.vW - 1);
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  const float inv_v = 1.0f / vscale;
  constexpr int NV = C::NV;
  for (int v = tid; v < NV; v += C::T) {
    const int lz = v / (BY * BX), ly = (v / BX) % BY, lx = v % BX;
    const int gz = lo0 + lz, gy = lo1 + ly, gx = lo2 + lx;
    if (gz > hi0 || gy > hi1 || gx > hi2) continue;
    const int o = ((lz + 1) * SY + (ly + 1)) * SX + (lx + 1);
    if (nseg > 1) {  // partial sums of a rotation segment: combined later
      u64 *dst = segbuf + ((int64_t)un.slot + seg) * (C::NACC * NV);
      dst[v] = s_tile[o];
      dst[NV + v] = s_tile[NT + o];
      if (MODE == kInsert) dst[2 * NV + v] = s_tile[2 * NT + o];
      continue;
    }
    f2v val = {(float)(long long)s_tile[o] * inv_v, (float)(long long)s_tile[NT + o] * inv_v};
    if (MODE == kInsert) {
      const float w = (float)(long long)s_tile[2 * NT + o] * (1.0f / kWScale);
      if (gx == 0) {  // the x = 0 plane goes through the Friedel mirror pass
        plane[(int64_t)gz * g.vH + gy] = val;
        wplane[(int64_t)gz * g.vH + gy] = w;
        continue;
      }
      const int64_t e = gz * sz + gy * sy + gx;
      wout[e] = accumulate ? wout[e] + w : w;
      out[e] = accumulate ? out[e] + val : val;
    } else {
      const int64_t e = gz * sz + gy * sy + gx;
      out[e] = accumulate ? out[e] + val : val;
    }
  }
}

// Sums the rotation segments of the split bricks (integer sums: the result
// does not depend on the order) and writes them like the brick flush.
template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__(256) void segment_combine_kernel(
    SliceGeom g, BrickGrid bg, const SplitRec *__restrict__ splits, const u64 *__restrict__ segbuf,
    ScaleArgs sa, int accumulate, f2v *__restrict__ out, float *__restrict__ wout,
    f2v *__restrict__ plane, float *__restrict__ wplane) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int NV = C::NV, NA = C::NACC;
  if (sa.guard && *sa.flag == 0u) return;
  const SplitRec sp = splits[blockIdx.x];
  const int bz = sp.brick / (bg.nby * bg.nbx);
  const int brem = sp.brick - bz * bg.nby * bg.nbx;
  const int by = brem / bg.nbx, bx = brem - by * bg.nbx;
  const int lo0 = bz * BZ, lo1 = by * BY, lo2 = bx * BX;
  float vscale, vlimit;
  scale_of(sa, vscale, vlimit);
  const float inv_v = 1.0f / vscale;
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  for (int v = threadIdx.x; v < NV; v += 256) {
    const int lz = v / (BY * BX), ly = (v / BX) % BY, lx = v % BX;
    const int gz = lo0 + lz, gy = lo1 + ly, gx = lo2 + lx;
    if (gz >= g.vD || gy >= g.vH || gx >= g.vW) continue;
    u64 acc[3] = {0ull, 0ull, 0ull};
    for (int sgi = 0; sgi < sp.nseg; ++sgi) {
      const u64 *b = segbuf + ((int64_t)sp.slot + sgi) * (NA * NV);
#pragma unroll
      for (int a = 0; a < NA; ++a) acc[a] += b[a * NV + v];
    }
    f2v val = {(float)(long long)acc[0] * inv_v, (float)(long long)acc[1] * inv_v};
    if (MODE == kInsert) {
      const float w = (float)(long long)acc[2] * (1.0f / kWScale);
      if (gx == 0) {
        plane[(int64_t)gz * g.vH + gy] = val;
        wplane[(int64_t)gz * g.vH + gy] = w;
        continue;
      }
      const int64_t e = gz * sz + gy * sy + gx;
      wout[e] = accumulate ? wout[e] + w : w;
      out[e] = accumulate ? out[e] + val : val;
    } else {
      const int64_t e = gz * sz + gy * sy + gx;
      out[e] = accumulate ? out[e] + val : val;
    }
  }
}

// max(|re|, |im|) of one 128-byte line in every 64 (or of every value when
// stride_lines == 1), as float bits into *res.
__global__ __launch_bounds__(256) void maxabs_lines_kernel(const f2v *__restrict__ x, int64_t n,
                                                           int64_t stride_lines, const unsigned *guard,
                                                           unsigned *__restrict__ res) {
  if (guard && *guard == 0u) return;
  const int64_t lanes = (int64_t)gridDim.x * 256;
  float m = 0.0f;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;; t += lanes) {
    const int64_t e = (t >> 4) * stride_lines * 16 + (t & 15);
    if (e >= n) break;
    const f2v v = x[e];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(res, __float_as_uint(m));
}

__global__ void set_word_kernel(unsigned *p, unsigned v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Gather zero pass: writes 0 to every pixel that is not certainly inside the
// volume and the disc.  One lane per slice row computes, for each Friedel
// half, the interval of columns whose canonical coordinates are inside the
// volume by a margin of 1e-7 voxel, whose half decision is certain and (with
// fftfreq_max) whose disc test is certain; the row's other columns (at most
// three runs) get a zero.  Runs before the brick gather, which overwrites
// the pixels of the margins that are inside after all.
// ---------------------------------------------------------------------------
struct ZeroRow {
  long long base;
  int a1, n1, a2, n2, a3, n3;  // three runs of columns [a, a + n)
};

__device__ __forceinline__ void lin_interval(double A, double B, double lo, double hi, double &jlo,
                                             double &jhi) {
  // lo <= A + j B <= hi
  if (fabs(B) > 1e-30) {
    const double t1 = (lo - A) / B, t2 = (hi - A) / B;
    jlo = fmax(jlo, fmin(t1, t2));
    jhi = fmin(jhi, fmax(t1, t2));
  } else if (A < lo || A > hi) {
    jhi = -1e300;
  }
}

__global__ __launch_bounds__(256) void gather_zero_kernel(SliceGeom g, const float *__restrict__ rot,
                                                          int64_t n_rows, f2v *__restrict__ out) {
  constexpr int CH = 4;
  __shared__ ZeroRow s_zr[4][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t t = ((int64_t)blockIdx.x * 4 + wave) * 64 + lane;
  int total = 0;
  if (t < n_rows) {
    const int64_t r = t / g.H;
    const int i = (int)(t - r * g.H);
    const Rot rr = load_rot(rot + r * 9, g);
    const double fy = canon_fy(g, i);
    const double m = 1e-7;
    const double top[3] = {(double)(g.vD - 1), (double)(g.vH - 1), (double)(g.vW - 1)};
    int a[2], b[2];
#pragma unroll
    for (int sh = 0; sh < 2; ++sh) {
      const double s = sh ? -1.0 : 1.0;
      double jlo = 0.0, jhi = (double)(g.Wr - 1);
      // certain Friedel half: the extraction decision flips within tolx of x' = 0
      const double ux = __dmul_rn(rr.U[2], fy);
      lin_interval(s * ux, s * rr.V[2], g.tolx + m, 1e300, jlo, jhi);
      const double C0 = __fma_rn(s * rr.U[0], fy, g.bz), C1 = __fma_rn(s * rr.U[1], fy, g.by);
      lin_interval(C0, s * rr.V[0], m, top[0] - m, jlo, jhi);
      lin_interval(C1, s * rr.V[1], m, top[1] - m, jlo, jhi);
      lin_interval(s * ux, s * rr.V[2], m, top[2] - m, jlo, jhi);
      if (g.use_mask) {
        // in_disc (float32) holds wherever the exact radius is 1e-6 below fmax
        const double fm = (double)g.fmax * (1.0 - 1e-6);
        const double fyv = (double)(i - g.half_H) * (double)g.inv_H32;
        const double rem = fm * fm - fyv * fyv;
        if (rem < 0.0 || fm < 0.0) jhi = -1e300;
        else jhi = fmin(jhi, sqrt(rem) / (double)g.inv_W32 - 1e-6);
      }
      a[sh] = 0;
      b[sh] = -1;
      if (jlo <= jhi) {
        a[sh] = max((int)ceil(jlo + 1e-7), 0);
        b[sh] = min((int)floor(jhi - 1e-7), g.Wr - 1);
      }
    }
    // sort the two certain intervals, drop empty ones, merge overlaps
    int a1 = a[0], b1 = b[0], a2 = a[1], b2 = b[1];
    if (a1 > b1) { a1 = a2; b1 = b2; a2 = 0; b2 = -1; }
    if (a2 <= b2 && a2 < a1) { int ta = a1, tb = b1; a1 = a2; b1 = b2; a2 = ta; b2 = tb; }
    if (a2 <= b2 && a2 <= b1 + 1) { b1 = max(b1, b2); a2 = 0; b2 = -1; }
    ZeroRow z;
    z.base = (long long)r * g.P + (long long)i * g.Wr;
    if (a1 > b1) {  // nothing certain: the whole row
      z.a1 = 0; z.n1 = g.Wr; z.a2 = 0; z.n2 = 0; z.a3 = 0; z.n3 = 0;
    } else if (a2 > b2) {
      z.a1 = 0; z.n1 = a1; z.a2 = b1 + 1; z.n2 = g.Wr - 1 - b1; z.a3 = 0; z.n3 = 0;
    } else {
      z.a1 = 0; z.n1 = a1; z.a2 = b1 + 1; z.n2 = a2 - b1 - 1; z.a3 = b2 + 1; z.n3 = g.Wr - 1 - b2;
    }
    total = z.n1 + z.n2 + z.n3;
    s_zr[wave][lane] = z;
  }
  const int nch = (total + CH - 1) / CH;
  const int cincl = wave_incl_scan(nch);
  const int total_ch = __shfl(cincl, 63, 64);
  __builtin_amdgcn_wave_barrier();
  for (int pc = 0; pc < total_ch; pc += 64) {
    const int pidx = pc + lane;
    const int k = find_segment<64>(cincl, pidx);
    const int cexcl_raw = __shfl(cincl, k > 0 ? k - 1 : 0, 64);
    const int cexcl = k > 0 ? cexcl_raw : 0;
    const bool live = pidx < total_ch;
    const ZeroRow z = s_zr[wave][k];
    const int q0 = (pidx - cexcl) * CH;
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int qq = q0 + q;
      int j;
      if (qq < z.n1) j = z.a1 + qq;
      else if (qq < z.n1 + z.n2) j = z.a2 + (qq - z.n1);
      else j = z.a3 + (qq - z.n1 - z.n2);
      if (live && qq < z.n1 + z.n2 + z.n3) out[z.base + j] = f2v{0.0f, 0.0f};
    }
  }
}

// x=0-plane Friedel mirror of the insertion, from the plane of direct splats
// (slice_insertion/_insert_central_slices_rfft_3d.py:350-404):
//   V[z,y,0] = P[z,y] + conj(P[mz,my])  unless (mz,my) is DC.
__global__ __launch_bounds__(256) void mirror_kernel(const f2v *__restrict__ plane,
                                                     const float *__restrict__ wplane, int D, int H,
                                                     int Wr, f2v *__restrict__ out,
                                                     float *__restrict__ wout, int accumulate) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
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

template <typename T>
__global__ __launch_bounds__(256) void fill_zero_kernel(T *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = (T)0;
}

template <typename T> int fill_zero(T *p, int64_t n, hipStream_t s) {
  if (n <= 0) return TT_OK;
  const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(fill_zero_kernel<T>, dim3(grid), dim3(256), 0, s, p, n);
  return check_launch("fill_zero_kernel");
}

// ---------------------------------------------------------------------------
// Unit planning (host) with a per-device cache of the unit tables
// ---------------------------------------------------------------------------
struct PlanKey {
  int dev, bz, by, bx, scatter, vD, vH, vW, H, Wr, imgD, nrot_q, mask_bits;
  bool operator<(const PlanKey &o) const {
    return std::tie(dev, bz, by, bx, scatter, vD, vH, vW, H, Wr, imgD, nrot_q, mask_bits) <
           std::tie(o.dev, o.bz, o.by, o.bx, o.scatter, o.vD, o.vH, o.vW, o.H, o.Wr, o.imgD, o.nrot_q,
                    o.mask_bits);
  }
};

struct PlanEntry {
  Unit *units = nullptr;
  int n_units = 0;
  SplitRec *splits = nullptr;
  int n_splits = 0;
  int n_slots = 0;  // rotation-segment partial slots of the split bricks
};

std::mutex g_plan_mutex;
std::map<PlanKey, PlanEntry> g_plans;

constexpr int kTargetSlots = 256 * 2 * 4;  // CUs x resident blocks x units per block slot

// Units of the bricks, heaviest first.  The work of a brick is estimated for
// uniformly oriented planes: a plane through DC crosses a brick of
// half-diagonal rb at distance d with probability ~ min(1, rb / d) and covers
// ~ (BZ BY BX)^(2/3) of its pixels.  Bricks no slice sample can reach are
// dropped (gather) or kept as empty units that write zeros (scatter).
int get_plan(const SliceGeom &g, int BZ, int BY, int BX, bool scatter, int64_t n_rot, PlanEntry &pe) {
  int dev = 0;
  hipGetDevice(&dev);
  int64_t nq = 64;
  while (nq < n_rot) nq <<= 1;
  float fm = g.use_mask ? g.fmax : -1.0f;
  int mbits;
  std::memcpy(&mbits, &fm, 4);
  const PlanKey key{dev, BZ, BY, BX, scatter ? 1 : 0, g.vD, g.vH, g.vW, g.H, g.Wr,
                    (int)std::lround(g.az * 1e3), (int)nq, mbits};
  std::lock_guard<std::mutex> lock(g_plan_mutex);
  auto it = g_plans.find(key);
  if (it != g_plans.end()) {
    pe = it->second;
    return TT_OK;
  }
  // gather bricks tile the cells 0..N-2, scatter bricks the voxels 0..N-1
  const int sz_ = scatter ? 0 : 1;
  const int nbz = (g.vD - sz_ + BZ - 1) / BZ, nby = (g.vH - sz_ + BY - 1) / BY,
            nbx = (g.vW - sz_ + BX - 1) / BX;
  // largest distance of a slice sample from DC (array units), plus margin
  const double amax = std::max(g.az, std::max(g.ay, g.ax));
  const double fy_max = 0.5 + 1.0 / g.H, fx_max = 0.5 + 1.0 / (2.0 * (g.Wr - 1) + 1.0);
  double rmax = amax * std::sqrt(fy_max * fy_max + fx_max * fx_max);
  if (g.use_mask) rmax = std::min(rmax, amax * (double)std::max(g.fmax, 0.0f) * 1.01);
  rmax += 2.0;
  const double rb = 0.5 * std::sqrt((double)(BZ * BZ + BY * BY + BX * BX)) + 1.0;
  const double area = 1.2 * std::pow((double)BZ * BY * BX, 2.0 / 3.0);
  struct Br {
    int b;
    double est;
  };
  std::vector<Br> br;
  std::vector<int> empty;
  double total = 0.0;
  const int Nv[3] = {g.vD, g.vH, g.vW};
  const int Bd[3] = {BZ, BY, BX};
  for (int z = 0; z < nbz; ++z)
    for (int y = 0; y < nby; ++y)
      for (int x = 0; x < nbx; ++x) {
        // coordinate box of the brick's candidate cells
        const int b3[3] = {z, y, x};
        const double c[3] = {g.bz, g.by, 0.0};
        double dmin2 = 0.0, dc2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          const int lo = b3[d] * Bd[d];
          const double clo = scatter ? std::max(lo - 1, 0) : lo;
          const double chi = std::min(lo + Bd[d] - 1, Nv[d] - 2) + 1.0;
          const double q = std::max(0.0, std::max(clo - c[d], c[d] - chi));
          dmin2 += q * q;
          const double m = 0.5 * (clo + chi) - c[d];
          dc2 += m * m;
        }
        const int bid = (z * nby + y) * nbx + x;
        if (std::sqrt(dmin2) > rmax) {  // no slice sample reaches this brick
          if (scatter) empty.push_back(bid);
          continue;
        }
        const double dc = std::sqrt(dc2);
        const double est = (double)nq * std::min(1.0, rb / std::max(dc, 1e-9)) * area;
        br.push_back({bid, est});
        total += est;
      }
  const double target = std::max(total / kTargetSlots, 8192.0);
  const int max_seg = (int)std::max<int64_t>(1, std::min<int64_t>(nq / 256, 64));
  struct U {
    Unit u;
    double est;
  };
  std::vector<U> us;
  std::vector<SplitRec> splits;
  int slots = 0;
  for (const Br &b : br) {
    const int nseg = std::max(1, std::min(max_seg, (int)std::ceil(b.est / target)));
    int slot = -1;
    if (nseg > 1 && scatter) {
      slot = slots;
      splits.push_back(SplitRec{b.b, nseg, slot});
      slots += nseg;
    }
    for (int sgi = 0; sgi < nseg; ++sgi)
      us.push_back({Unit{b.b, (sgi << 16) | nseg, slot}, b.est / nseg});
  }
  for (int b : empty) us.push_back({Unit{b, 0, -1}, 0.0});
  std::stable_sort(us.begin(), us.end(), [](const U &a, const U &b) { return a.est > b.est; });
  std::vector<Unit> host(us.size());
  for (size_t k = 0; k < us.size(); ++k) host[k] = us[k].u;
  PlanEntry e;
  e.n_units = (int)host.size();
  e.n_splits = (int)splits.size();
  e.n_slots = slots;
  hipError_t err = hipSuccess;
  if (e.n_units > 0) {
    err = hipMalloc((void **)&e.units, host.size() * sizeof(Unit));
    if (err == hipSuccess)
      err = hipMemcpy(e.units, host.data(), host.size() * sizeof(Unit), hipMemcpyHostToDevice);
  }
  if (err == hipSuccess && e.n_splits > 0) {
    err = hipMalloc((void **)&e.splits, splits.size() * sizeof(SplitRec));
    if (err == hipSuccess)
      err = hipMemcpy(e.splits, splits.data(), splits.size() * sizeof(SplitRec), hipMemcpyHostToDevice);
  }
  if (err != hipSuccess) {
    set_error(std::string("brick plan upload: ") + hipGetErrorString(err));
    return TT_ERR_HIP;
  }
  g_plans[key] = e;
  pe = e;
  return TT_OK;
}

struct Scratch {
  void *p = nullptr;
  hipStream_t s;
  explicit Scratch(hipStream_t st) : s(st) {}
  int alloc(size_t bytes) {
    hipError_t e = hipMallocAsync(&p, bytes, s);
    if (e != hipSuccess) {
      p = nullptr;
      set_error(std::string("hipMallocAsync: ") + hipGetErrorString(e));
      return TT_ERR_HIP;
    }
    return TT_OK;
  }
  ~Scratch() {
    if (p) hipFreeAsync(p, s);
  }
};

constexpr int kGBZ = 16, kGBY = 16, kGBX = 8;  // gather brick (cells)
constexpr int kABZ = 16, kABY = 16, kABX = 8;  // adjoint brick (owned voxels)
constexpr int kIBZ = 16, kIBY = 8, kIBX = 8;   // insertion brick (owned voxels)

BrickGrid grid_of(const SliceGeom &g, int BZ, int BY, int BX, bool scatter) {
  const int sz_ = scatter ? 0 : 1;
  BrickGrid bg;
  bg.nbz = (g.vD - sz_ + BZ - 1) / BZ;
  bg.nby = (g.vH - sz_ + BY - 1) / BY;
  bg.nbx = (g.vW - sz_ + BX - 1) / BX;
  return bg;
}

int check_engine(const SliceGeom &g, int64_t n_rot) {
  if (n_rot >= ((int64_t)1 << 30)) return fail_invalid("too many rotations for one call (< 2^30)");
  if (g.H >= 32768) return fail_invalid("slice height must be < 32768");
  return TT_OK;
}

// Scatter driver shared by the adjoint and the insertion.
template <int BZ, int BY, int BX, int MODE>
int run_scatter(const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                float *wout, int accumulate, hipStream_t s) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  PlanEntry pe;
  if (int st = get_plan(g, BZ, BY, BX, true, n_rot, pe)) return st;
  const BrickGrid bg = grid_of(g, BZ, BY, BX, true);
  const size_t np = (size_t)g.vD * g.vH;
  // scratch: words (max bits sampled, flag, max bits exact) | plane | wplane | PlaneRec | segments
  size_t off = 256;
  const size_t off_plane = off;
  if (MODE == kInsert) off += np * sizeof(f2v);
  const size_t off_wplane = off;
  if (MODE == kInsert) off += np * sizeof(float);
  const size_t zero_end = off;
  off = (off + 255) & ~(size_t)255;
  const size_t off_pre = off;
  off += (size_t)std::max<int64_t>(n_rot, 1) * sizeof(PlaneRec);
  off = (off + 255) & ~(size_t)255;
  const size_t off_seg = off;
  off += (size_t)pe.n_slots * C::NACC * C::NV * sizeof(u64);
  Scratch sc(s);
  if (int st = sc.alloc(off)) return st;
  char *base = (char *)sc.p;
  unsigned *words = (unsigned *)base;
  f2v *plane = MODE == kInsert ? (f2v *)(base + off_plane) : nullptr;
  float *wplane = MODE == kInsert ? (float *)(base + off_wplane) : nullptr;
  PlaneRec *pre = (PlaneRec *)(base + off_pre);
  u64 *segbuf = (u64 *)(base + off_seg);
  if (int st = fill_zero<unsigned>(words, (int64_t)(zero_end / sizeof(unsigned)), s)) return st;
  const int64_t n = n_rot * (int64_t)g.P;
  // accumulate: exact max (the outputs cannot be rewritten); otherwise a
  // sampled max with 8 bits of headroom and an exact redo if it was exceeded
  const bool sampled = !accumulate;
  const int mgrid = (int)std::min<int64_t>(std::max<int64_t>(n / 4096, 1), 4096);
  if (n_rot > 0) {
    hipLaunchKernelGGL(maxabs_lines_kernel, dim3(mgrid), dim3(256), 0, s, src, n,
                       (int64_t)(sampled ? 64 : 1), (const unsigned *)nullptr, words);
    hipLaunchKernelGGL(plane_rec_kernel, dim3((unsigned)((n_rot + 255) / 256)), dim3(256), 0, s, g,
                       rot, (int)n_rot, pre);
    if (int st = check_launch("maxabs_lines / plane_rec")) return st;
  }
  for (int pass = 0; pass < (sampled ? 2 : 1); ++pass) {
    ScaleArgs sa;
    sa.maxbits = words + (pass ? 2 : 0);
    sa.flag = words + 1;
    sa.headroom = (sampled && pass == 0) ? 8 : 0;
    sa.guard = pass;
    if (pass == 1 && n_rot > 0) {
      hipLaunchKernelGGL(maxabs_lines_kernel, dim3(mgrid), dim3(256), 0, s, src, n, (int64_t)1,
                         (const unsigned *)(words + 1), words + 2);
      if (int st = check_launch("maxabs_lines_kernel")) return st;
    }
    if (pe.n_units > 0) {
      hipLaunchKernelGGL((brick_kernel<BZ, BY, BX, MODE>), dim3((unsigned)pe.n_units), dim3(C::T), 0,
                         s, src, g, bg, (const PlaneRec *)pre, (const Unit *)pe.units, (int)n_rot, out,
                         wout, plane, wplane, segbuf, sa, accumulate);
      if (int st = check_launch("brick_kernel")) return st;
    }
    if (pe.n_splits > 0) {
      hipLaunchKernelGGL((segment_combine_kernel<BZ, BY, BX, MODE>), dim3((unsigned)pe.n_splits),
                         dim3(256), 0, s, g, bg, (const SplitRec *)pe.splits, (const u64 *)segbuf, sa,
                         accumulate, out, wout, plane, wplane);
      if (int st = check_launch("segment_combine_kernel")) return st;
    }
  }
  if (MODE == kInsert) {
    hipLaunchKernelGGL(mirror_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, plane, wplane,
                       g.vD, g.vH, g.vW, out, wout, accumulate);
    if (int st = check_launch("mirror_kernel")) return st;
  }
  return TT_OK;
}

}  // namespace

int engine_extract(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                   hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0) return TT_OK;
  const int64_t n_rows = n_rot * g.H;
  const int64_t nblk = (n_rows + 255) / 256;
  if (nblk >= ((int64_t)1 << 31)) return fail_invalid("too many slice rows for one call");
  hipLaunchKernelGGL(gather_zero_kernel, dim3((unsigned)nblk), dim3(256), 0, s, g, rot, n_rows, out);
  if (int st = check_launch("gather_zero_kernel")) return st;
  PlanEntry pe;
  if (int st = get_plan(g, kGBZ, kGBY, kGBX, false, n_rot, pe)) return st;
  if (pe.n_units == 0) return TT_OK;
  Scratch sc(s);
  if (int st = sc.alloc((size_t)n_rot * sizeof(PlaneRec))) return st;
  PlaneRec *pre = (PlaneRec *)sc.p;
  hipLaunchKernelGGL(plane_rec_kernel, dim3((unsigned)((n_rot + 255) / 256)), dim3(256), 0, s, g, rot,
                     (int)n_rot, pre);
  if (int st = check_launch("plane_rec_kernel")) return st;
  const BrickGrid bg = grid_of(g, kGBZ, kGBY, kGBX, false);
  ScaleArgs sa = {nullptr, nullptr, 0, 0};
  hipLaunchKernelGGL((brick_kernel<kGBZ, kGBY, kGBX, kGather>), dim3((unsigned)pe.n_units),
                     dim3(Cfg<kGBZ, kGBY, kGBX, kGather>::T), 0, s, vol, g, bg, (const PlaneRec *)pre,
                     (const Unit *)pe.units, (int)n_rot, out, (float *)nullptr, (f2v *)nullptr,
                     (float *)nullptr, (u64 *)nullptr, sa, 0);
  return check_launch("brick_kernel");
}

int engine_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot, f2v *gV,
                   int accumulate, hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0 && accumulate) return TT_OK;
  return run_scatter<kABZ, kABY, kABX, kAdjoint>(g, gS, rot, n_rot, gV, nullptr, accumulate, s);
}

int engine_insert(const SliceGeom &g, const f2v *slices, const float *rot, int64_t n_rot, f2v *vol,
                  float *w, int accumulate, hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0 && accumulate) return TT_OK;
  return run_scatter<kIBZ, kIBY, kIBX, kInsert>(g, slices, rot, n_rot, vol, w, accumulate, s);
}

}  // namespace tt
