// Brick-stationary central-slice gather (extraction) for gfx950.
//
// The pixel-major gather reads each sample's 8 corners through L2; a plane of
// arbitrary orientation uses few voxels of each 128-byte line, and the 67.6 MB
// volume of a 256^3 rfft does not fit the 4 MB L2 of an XCD, so nearly every
// corner read misses.  Here a workgroup owns a brick of trilinear cells, copies
// its voxels (cells + 1 halo voxel per axis) into LDS once, and then serves
// every slice crossing the brick from LDS:
//  (1) one lane per rotation tests the slice plane against the brick and, for
//      each Friedel half that reaches it, records the range of slice rows;
//  (2) one wave per (rotation, half) item: one lane per row computes the row's
//      candidate column interval from three slab constraints (FP64, widened by
//      1e-6 voxel); a wave prefix sum spreads the candidate pixels over the
//      lanes, consecutive lanes taking consecutive pixels of a row;
//  (3) each candidate pixel is mapped with the exact coordinate function of
//      the pixel-major gather (FP64, integer snap, tolerance conjugation); the
//      pixel is written by the brick that holds its clamped lower-corner cell,
//      so every pixel inside the volume is written exactly once.
// Pixels outside the volume or outside the fftfreq_max disc belong to no
// brick; a light pixel-major pass writes their zeros.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_gather.h"

namespace tt {

namespace {

constexpr int kGT = 256;
constexpr int kGWaves = kGT / 64;
constexpr int kGRotChunk = kGT;
constexpr int kGItemsMax = 2 * kGRotChunk;
constexpr double kGMargin = 1e-6;
constexpr int kGMaxSeg = 8;

struct GatherGrid {
  int nbz, nby, nbx, n_bricks;  // bricks of cells; cells run 0..N-2 per axis
  int smax;                     // units per brick (rotation segments, upper bound)
  int rot_per_unit;             // target crossing rotations per unit
  double rb;                    // brick half-diagonal (voxels)
};

__device__ __forceinline__ int uniform_int(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Rotation segments of a brick: the fraction of uniformly oriented central
// planes passing within rb of a point at distance d is rb/d.
__device__ __forceinline__ int brick_segments(const SliceGeom &g, const GatherGrid &gg, int n_rot,
                                              double cz, double cy, double cx) {
  const double dz = cz - g.bz, dy = cy - g.by;
  const double d = sqrt(dz * dz + dy * dy + cx * cx);
  const double frac = d > gg.rb ? gg.rb / d : 1.0;
  const int est = (int)ceil(frac * (double)n_rot);
  return min(max((est + gg.rot_per_unit - 1) / gg.rot_per_unit, 1), gg.smax);
}

template <int BZ, int BY, int BX>
__global__ __launch_bounds__(kGT) void brick_gather_kernel(const f2v *__restrict__ vol, SliceGeom g,
                                                           const float *__restrict__ rot, int n_rot,
                                                           GatherGrid gg, f2v *__restrict__ out) {
  constexpr int SX = BX + 1, SY = BY + 1, SZ = BZ + 1;
  constexpr int NT = SZ * SY * SX;
  __shared__ f2v s_vol[NT];
  __shared__ int s_code[kGItemsMax];  // 2 * rotation + (half == -1)
  __shared__ int s_rows[kGItemsMax];  // first row | (last row << 16), rows as m = i - H/2
  __shared__ int s_nitems;

  // Units run segment-major, highest segment first: the dispatcher deals
  // consecutive blocks to the 8 XCDs in turn, so consecutive bricks (of
  // similar load) spread over all XCDs, and the extra segments of the heavy
  // bricks near DC start early instead of forming the tail.
  const int unit = blockIdx.x;
  const int layer = unit / gg.n_bricks;
  const int brick = unit - layer * gg.n_bricks, seg = gg.smax - 1 - layer;
  if (seg < 0) return;
  const int bz = brick / (gg.nby * gg.nbx);
  const int rem = brick - bz * gg.nby * gg.nbx;
  const int by = rem / gg.nbx, bx = rem - by * gg.nbx;
  const int lo[3] = {bz * BZ, by * BY, bx * BX};
  const int hc[3] = {min(lo[0] + BZ - 1, g.vD - 2), min(lo[1] + BY - 1, g.vH - 2),
                     min(lo[2] + BX - 1, g.vW - 2)};  // last cell of the brick
  const int nseg = brick_segments(g, gg, n_rot, 0.5 * (lo[0] + hc[0] + 1),
                                  0.5 * (lo[1] + hc[1] + 1), 0.5 * (lo[2] + hc[2] + 1));
  if (seg >= nseg) return;
  const int r_begin = (int)((int64_t)n_rot * seg / nseg);
  const int r_end = (int)((int64_t)n_rot * (seg + 1) / nseg);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  // voxel tile: cells lo..hc plus the upper corner of the last cell
  for (int v = tid; v < NT; v += kGT) {
    const int z = v / (SY * SX), y = (v / SX) % SY, x = v % SX;
    const int gz = lo[0] + z, gy = lo[1] + y, gx = lo[2] + x;
    f2v val = {0.0f, 0.0f};
    if (gz < g.vD && gy < g.vH && gx < g.vW) val = vol[((int64_t)gz * g.vH + gy) * g.vW + gx];
    s_vol[v] = val;
  }

  double Lb[3], Ub[3], ctr[3], half[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Lb[d] = (double)lo[d] - kGMargin;
    Ub[d] = (double)(hc[d] + 1) + kGMargin;
    ctr[d] = 0.5 * (Lb[d] + Ub[d]);
    half[d] = 0.5 * (Ub[d] - Lb[d]);
  }
  const double beta[3] = {g.bz, g.by, 0.0};
  const double scale[3] = {g.az, g.ay, g.ax};
  const double mmin = (double)(-g.half_H), mmax = (double)(g.H - 1 - g.half_H);

  for (int r0 = r_begin; r0 < r_end; r0 += kGRotChunk) {
    if (tid == 0) s_nitems = 0;
    __syncthreads();
    // (1) plane / brick test, one lane per rotation
    const int r = r0 + tid;
    if (r < r_end) {
      const float *R = rot + (int64_t)r * 9;
      double A[3], B[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        A[d] = scale[d] * (double)R[3 * d + 1] * g.inv_H;
        B[d] = scale[d] * (double)R[3 * d + 2] * g.inv_W;
      }
      const double n0 = A[1] * B[2] - A[2] * B[1];
      const double n1 = A[2] * B[0] - A[0] * B[2];
      const double n2 = A[0] * B[1] - A[1] * B[0];
      const double dist = n0 * (ctr[0] - beta[0]) + n1 * (ctr[1] - beta[1]) + n2 * (ctr[2] - beta[2]);
      const double rad = fabs(n0) * half[0] + fabs(n1) * half[1] + fabs(n2) * half[2];
      const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      if (fabs(dist) <= rad + 1e-9 * nn) {
        const double AA = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
        const double BB = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
        const double AB = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
        const double det = AA * BB - AB * AB;
        double mc = 0.0, mr = 0.0, jc = 0.0, jr = 0.0;
        if (det > 1e-20 * AA * BB) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double pm = (BB * A[d] - AB * B[d]) / det;
            const double pj = (AA * B[d] - AB * A[d]) / det;
            mc += pm * (ctr[d] - beta[d]);
            jc += pj * (ctr[d] - beta[d]);
            mr += fabs(pm) * half[d];
            jr += fabs(pj) * half[d];
          }
          mr += 1e-6;
          jr += 1e-6;
        } else {  // degenerate matrix: every row and column is a candidate
          mr = jr = 1e300;
        }
        for (int sh = 0; sh < 2; ++sh) {
          const double s = sh ? -1.0 : 1.0;
          if (s * jc + jr < 0.0 || s * jc - jr > (double)(g.Wr - 1)) continue;
          const double mlo = fmax(s * mc - mr, mmin), mhi = fmin(s * mc + mr, mmax);
          const int m0 = (int)ceil(mlo), m1 = (int)floor(mhi);
          if (m0 > m1) continue;
          const int slot = atomicAdd(&s_nitems, 1);
          s_code[slot] = 2 * r + sh;
          s_rows[slot] = (int)(((unsigned)m1 << 16) | ((unsigned)m0 & 0xffffu));
        }
      }
    }
    __syncthreads();
    const int nitems = s_nitems;
    // (2) one wave per item
    for (int it = wave; it < nitems; it += kGWaves) {
      const int code = uniform_int(s_code[it]);
      const int rows = uniform_int(s_rows[it]);
      const int m0 = (int)(short)(rows & 0xffff), m1 = rows >> 16;
      const int ri = code >> 1;
      const bool neg = code & 1;
      const double s = neg ? -1.0 : 1.0;
      const float *R = rot + (int64_t)ri * 9;
      const Rot rr = load_rot(R, g);
      double sA[3], sB[3], inv[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        sA[d] = s * scale[d] * (double)R[3 * d + 1] * g.inv_H;
        sB[d] = s * scale[d] * (double)R[3 * d + 2] * g.inv_W;
        inv[d] = fabs(sB[d]) > 1e-12 ? 1.0 / sB[d] : 0.0;
      }
      f2v *orow = out + (int64_t)ri * g.P;
      for (int mb = m0; mb <= m1; mb += 64) {
        const int m = mb + lane;
        int ja = 0, len = 0;
        if (m <= m1) {
          double jmin = 0.0, jmax = (double)(g.Wr - 1);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double base = beta[d] + sA[d] * (double)m;
            if (inv[d] != 0.0) {
              const double t1 = (Lb[d] - base) * inv[d], t2 = (Ub[d] - base) * inv[d];
              jmin = fmax(jmin, fmin(t1, t2));
              jmax = fmin(jmax, fmax(t1, t2));
            } else if (base < Lb[d] || base > Ub[d]) {
              jmax = -1.0;
            }
          }
          // [jmin, jmax] holds every column whose exact coordinate can lie in
          // the brick: the box is widened by 1e-6 voxel, far above the FP64
          // rounding of this estimate; the exact test below decides.
          ja = (int)ceil(jmin);
          len = max(0, (int)floor(jmax) - ja + 1);
        }
        int incl = len;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int t = __shfl_up(incl, off, 64);
          if (lane >= off) incl += t;
        }
        const int total = uniform_int(__shfl(incl, 63, 64));
        const int excl = incl - len;
        for (int p0 = 0; p0 < total; p0 += 64) {
          const int p = p0 + lane;
          // row k: first lane whose inclusive prefix exceeds p
          int k = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1) {
            const int v = __shfl(incl, k + step - 1, 64);
            if (v <= p) k += step;
          }
          const int k_ja = __shfl(ja, k, 64);
          const int k_ex = __shfl(excl, k, 64);
          if (p >= total) continue;
          const int j = k_ja + (p - k_ex);
          const int i = mb + k + g.half_H;
          if (!in_disc(g, i, j)) continue;
          const Coord c = slice_coord<false>(g, rr, i, j);
          if (c.conj != neg || !inside(g, c)) continue;
          const Cell ce = cell_of(g, c);
          const int z0 = ce.z0 - lo[0], y0 = ce.y0 - lo[1], x0 = ce.x0 - lo[2];
          if (z0 < 0 || z0 >= BZ || y0 < 0 || y0 >= BY || x0 < 0 || x0 >= BX) continue;
          const int e = (z0 * SY + y0) * SX + x0;
          const f2v v000 = s_vol[e], v001 = s_vol[e + 1];
          const f2v v010 = s_vol[e + SX], v011 = s_vol[e + SX + 1];
          const f2v v100 = s_vol[e + SY * SX], v101 = s_vol[e + SY * SX + 1];
          const f2v v110 = s_vol[e + SY * SX + SX], v111 = s_vol[e + SY * SX + SX + 1];
          // same weights and order as the pixel-major gather: x, then y, then z
          const float ux = 1.0f - ce.tx, uy = 1.0f - ce.ty, uz = 1.0f - ce.tz;
          const float r00 = v000.x * ux + v001.x * ce.tx, i00 = v000.y * ux + v001.y * ce.tx;
          const float r01 = v010.x * ux + v011.x * ce.tx, i01 = v010.y * ux + v011.y * ce.tx;
          const float r10 = v100.x * ux + v101.x * ce.tx, i10 = v100.y * ux + v101.y * ce.tx;
          const float r11 = v110.x * ux + v111.x * ce.tx, i11 = v110.y * ux + v111.y * ce.tx;
          const float r0 = r00 * uy + r01 * ce.ty, i0 = i00 * uy + i01 * ce.ty;
          const float r1 = r10 * uy + r11 * ce.ty, i1 = i10 * uy + i11 * ce.ty;
          f2v val = {r0 * uz + r1 * ce.tz, i0 * uz + i1 * ce.tz};
          if (c.conj) val.y = -val.y;
          orow[(int64_t)i * g.Wr + j] = val;
        }
      }
    }
    __syncthreads();
  }
}

// Zeros for the pixels no brick writes: outside the volume or the disc.  An
// FP32 estimate settles clear cases; pixels within 1e-3 voxel of a face (or
// near x' = 0) take the exact coordinate function.
__global__ __launch_bounds__(kGT) void gather_zero_kernel(SliceGeom g, const float *__restrict__ rot,
                                                          int64_t blocks_per_rot, int64_t n_blocks,
                                                          f2v *__restrict__ out) {
  const int64_t b = blockIdx.x;
  if (b >= n_blocks) return;
  const int64_t r = b / blocks_per_rot;
  const int p = (int)((b - r * blocks_per_rot) * kGT) + threadIdx.x;
  if (p >= g.P) return;
  int i, j;
  split_pixel(g, p, i, j);
  const float *R = rot + r * 9;
  bool zero = !in_disc(g, i, j);
  if (!zero) {
    const float fy = (float)(i - g.half_H) * (float)g.inv_H, fx = (float)j * (float)g.inv_W;
    const float qz = R[1] * fy + R[2] * fx, qy = R[4] * fy + R[5] * fx, qx = R[7] * fy + R[8] * fx;
    const float cx = fabsf(qx) * (float)g.ax;
    const float sgn = qx < 0.0f ? -1.0f : 1.0f;
    const float cz = sgn * qz * (float)g.az + (float)g.bz, cy = sgn * qy * (float)g.ay + (float)g.by;
    // FP32 error of these coordinates: a few 2^-24 relative of q, times the
    // array scale; e is well above it.
    const float e = 1e-3f + 2e-6f * (float)fmax(fmax(g.az, g.ay), g.ax);
    const float tz = (float)(g.vD - 1), ty = (float)(g.vH - 1), tx = (float)(g.vW - 1);
    // near x' = 0 the Friedel half (which flips cz and cy) is decided exactly
    const bool half_known = cx > e;
    const bool clear_in = half_known && cx < tx - e && cz > e && cz < tz - e && cy > e && cy < ty - e;
    const bool clear_out =
        half_known && (cx > tx + e || cz < -e || cz > tz + e || cy < -e || cy > ty + e);
    if (clear_out) {
      zero = true;
    } else if (!clear_in) {
      const Rot rr = load_rot(R, g);
      zero = !inside(g, slice_coord<false>(g, rr, i, j));
    }
  }
  if (zero) out[r * g.P + p] = f2v{0.0f, 0.0f};
}

template <int BZ, int BY, int BX>
int launch_brick_gather(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot,
                        f2v *out, hipStream_t s) {
  GatherGrid gg;
  gg.nbz = (g.vD - 1 + BZ - 1) / BZ;
  gg.nby = (g.vH - 1 + BY - 1) / BY;
  gg.nbx = (g.vW - 1 + BX - 1) / BX;
  gg.n_bricks = gg.nbz * gg.nby * gg.nbx;
  gg.smax = kGMaxSeg;
  gg.rot_per_unit = 512;
  gg.rb = 0.5 * std::sqrt((double)(BZ * BZ + BY * BY + BX * BX)) + 1.0;
  const int64_t bpr = (g.P + kGT - 1) / kGT;
  const int64_t kMaxBlocks = (int64_t)1 << 23;
  for (int64_t r0 = 0; r0 < n_rot; r0 += std::max<int64_t>(1, kMaxBlocks / bpr)) {
    const int64_t nr = std::min<int64_t>(std::max<int64_t>(1, kMaxBlocks / bpr), n_rot - r0);
    hipLaunchKernelGGL(gather_zero_kernel, dim3((unsigned)(nr * bpr)), dim3(kGT), 0, s, g,
                       rot + r0 * 9, bpr, nr * bpr, out + r0 * g.P);
    if (int st = check_launch("gather_zero_kernel")) return st;
  }
  // rotations per launch bounded so a row code (2 r + 1) stays in an int
  const int64_t kMaxRot = (int64_t)1 << 29;
  for (int64_t r0 = 0; r0 < n_rot; r0 += kMaxRot) {
    const int nr = (int)std::min<int64_t>(kMaxRot, n_rot - r0);
    hipLaunchKernelGGL((brick_gather_kernel<BZ, BY, BX>), dim3((unsigned)(gg.n_bricks * gg.smax)),
                       dim3(kGT), 0, s, vol, g, rot + r0 * 9, nr, gg, out + r0 * g.P);
    if (int st = check_launch("brick_gather_kernel")) return st;
  }
  return TT_OK;
}

}  // namespace

bool brick_gather_eligible(const SliceGeom &g, int64_t /*n_rot*/) {
  // The brick gather needs cells on every axis and 16-bit row indices.  It is
  // not chosen automatically yet: at 256^3 x 4096 it spends about 200 vector
  // instructions per 64-pixel pass on exact per-pixel coordinates and row
  // decoding, and runs at 2.3 ms against 1.8 ms for the pixel-major gather.
  return g.vD >= 2 && g.vH >= 2 && g.vW >= 2 && g.H <= 32767;
}

int brick_gather(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                 hipStream_t s) {
  return launch_brick_gather<16, 16, 8>(g, vol, rot, n_rot, out, s);
}

}  // namespace tt
