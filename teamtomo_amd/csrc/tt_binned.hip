// Binned scatter: the extraction adjoint and the insertion when the slice
// pixels are few per volume brick, as for a tilt series inserted into a large
// volume (61 slices of 512^2 into 512^3: ~300 pixels per 8x16x16 brick).
//
// The brick engine discovers each brick's slices itself (plane tests, FP64
// item frames, row and chunk lists), a chain of dependent steps that costs
// more than the few pixels it then scatters; its upper-shell halos also make
// a second pass read and rewrite every cache line of the volume.  Here the
// pixels come to the bricks instead:
//   1. bin_count_kernel: one lane per pixel evaluates the canonical FP64
//      coordinate (the pixel-atomic kernels' code, tt_common.h) and counts it
//      in every brick that owns one of its 8 trilinear corners with a
//      nonzero weight (1, 2, 4 or 8 bricks; wave-aggregated atomics);
//   2. bin_scan_kernel: exclusive prefix of the counts;
//   3. bin_fill_kernel: the same evaluation again, each pixel written as a
//      24-byte record (corner offset, fractions, value) into its bricks' bins;
//   4. bin_brick_kernel: one block per brick sums its records' owned corners
//      in LDS and writes every voxel it owns once (no halo pass), zeros for an
//      empty brick;
//   5. the insertion's x = 0 Friedel mirror pass (shared with the engine).
//
// Accumulation is 64-bit fixed point at a per-brick scale from the exact
// maximum of the brick's records: each FP32 product w v is converted exactly
// (to_fixed64), so the sums are exact and independent of the record order
// (which the atomics of pass 3 leave unspecified): the result is bitwise
// reproducible.  Semantics as the pixel-atomic kernels (tt_scatter.hip), i.e.
//   insertion  slice_insertion/_insert_central_slices_rfft_3d.py:9-106,213-406
//   adjoint    the transpose of slice_extraction/_extract_central_slices_rfft_3d.py:84-222
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_engine.h"
#include "tt_fixed.h"

namespace tt {

namespace {

constexpr int kBZ = 8, kBY = 16, kBX = 16;  // voxels a brick owns per axis
constexpr int kNV = kBZ * kBY * kBX;
constexpr int kThreads = 256;
// A pixel lands in at most 8 bricks: the record buffer holds 8 per pixel, and
// the path is used only while that buffer stays below this size.
constexpr int64_t kMaxRecordBytes = (int64_t)3 << 30;
// Auto selection: at most this many pixels (slices x pixels per slice) per
// brick of the volume on average.  Measured (insertion / adjoint, ms, brick
// engine against binned): 61 tilts of 512^2 into 512^3 (230 per brick)
// 2.29 / 1.45 against 0.99 / 0.80; 61 of 256^2 into 256^3 (437) 0.39 / 0.27
// against 0.41 / 0.39; 256 random slices into 256^3 (1833) 0.48 / 0.31
// against 1.14 / 1.07 (the bins of the bricks near DC then take most records,
// and their counters serialize the atomics).
constexpr double kAutoPixelsPerBrick = 320.0;

struct __align__(8) BinRec {
  unsigned off;  // corner offsets o + 1 (o in [-1, B-1]) per axis: z | y << 8 | x << 16
  float t[3];    // fractions (z, y, x) of the pixel in its cell
  float v[2];    // the value, conjugated in the lower Friedel half
};
static_assert(sizeof(BinRec) == 24, "24-byte bin records");

struct BinGrid {
  int nbz, nby, nbx;
};

int g_scatter_path = TT_SCATTER_AUTO;

__device__ __forceinline__ int lane_of() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ int rank_below_mask(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Adds 1 to ctr[key] for every lane with key >= 0, one atomic per distinct
// key in the wave; with RET, returns the lane's slot (the counter's value
// before its own increment, lanes of one key taking consecutive slots).
// Called with the whole wave active.
template <bool RET>
__device__ __forceinline__ int wave_claim(int key, int *__restrict__ ctr) {
  const int lane = lane_of();
  int slot = -1;
  unsigned long long pending = __ballot(key >= 0);
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int k = __builtin_amdgcn_readlane(key, leader);
    const unsigned long long m = __ballot(key == k);
    if (RET) {
      int base = 0;
      if (lane == leader) base = atomicAdd(ctr + k, __popcll(m));
      base = __builtin_amdgcn_readlane(base, leader);
      if (key == k) slot = base + rank_below_mask(m);
    } else if (lane == leader) {
      atomicAdd(ctr + k, __popcll(m));
    }
    pending &= ~m;
  }
  return slot;
}

// A slice pixel's cell, as the pixel-atomic scatter evaluates it; false for a
// pixel the scatter skips (past the slice, the insertion's x = 0, y < 0 half
// row, outside the disc or the volume).
template <bool INSERT>
__device__ __forceinline__ bool bin_pixel(const SliceGeom &g, const Rot &r, int p, Cell &k, bool &conj) {
  if (p >= g.P) return false;
  int i, j;
  split_pixel(g, p, i, j);
  if (INSERT && j == 0 && i < g.half_H) return false;
  const Coord c = slice_coord<INSERT>(g, r, i, j);
  if (!(in_disc(g, i, j) && inside(g, c))) return false;
  k = cell_of(g, c);
  conj = c.conj;
  return true;
}

// The bricks a cell reaches on each axis: the lower corner's brick if its
// weight 1 - t is nonzero (or the upper corner shares it), the upper corner's
// if its weight t is nonzero and it lies in the next brick.
struct Targets {
  int blo[3];
  unsigned lo_ok, hi_ok;  // bit d: axis d
};

__device__ __forceinline__ Targets targets_of(const Cell &k) {
  Targets tg;
  const int c[3] = {k.z0, k.y0, k.x0};
  const float t[3] = {k.tz, k.ty, k.tx};
  const int B[3] = {kBZ, kBY, kBX};
  tg.lo_ok = tg.hi_ok = 0u;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int lo = c[d] / B[d], hi = (c[d] + 1) / B[d];
    tg.blo[d] = lo;
    if (t[d] < 1.0f || hi == lo) tg.lo_ok |= 1u << d;
    if (hi != lo && t[d] > 0.0f) tg.hi_ok |= 1u << d;
  }
  return tg;
}

// Brick of target q (bit 2: upper on z, bit 1: on y, bit 0: on x), or -1.
__device__ __forceinline__ int target_brick(const Targets &tg, int q, const BinGrid &bg) {
  const int cz = (q >> 2) & 1, cy = (q >> 1) & 1, cx = q & 1;
  const bool ok = ((cz ? tg.hi_ok : tg.lo_ok) >> 0 & 1u) && ((cy ? tg.hi_ok : tg.lo_ok) >> 1 & 1u) &&
                  ((cx ? tg.hi_ok : tg.lo_ok) >> 2 & 1u);
  if (!ok) return -1;
  return ((tg.blo[0] + cz) * bg.nby + (tg.blo[1] + cy)) * bg.nbx + (tg.blo[2] + cx);
}

template <bool INSERT>
__global__ __launch_bounds__(kThreads) void bin_count_kernel(SliceGeom g, const float *__restrict__ rot,
                                                            BinGrid bg, int *__restrict__ count) {
  const int rr = blockIdx.y;
  const Rot r = load_rot(rot + (int64_t)rr * 9, g);
  const int p = blockIdx.x * kThreads + threadIdx.x;
  Cell k;
  bool conj = false;
  const bool ok = bin_pixel<INSERT>(g, r, p, k, conj);
  Targets tg;
  if (ok) tg = targets_of(k);
#pragma unroll 1
  for (int q = 0; q < 8; ++q) wave_claim<false>(ok ? target_brick(tg, q, bg) : -1, count);
}

template <bool INSERT>
__global__ __launch_bounds__(kThreads) void bin_fill_kernel(SliceGeom g, const f2v *__restrict__ src,
                                                           const float *__restrict__ rot, BinGrid bg,
                                                           int *__restrict__ cursor, BinRec *__restrict__ recs,
                                                           int64_t cap) {
  const int rr = blockIdx.y;
  const Rot r = load_rot(rot + (int64_t)rr * 9, g);
  const int p = blockIdx.x * kThreads + threadIdx.x;
  Cell k;
  bool conj = false;
  const bool ok = bin_pixel<INSERT>(g, r, p, k, conj);
  f2v v = {0.0f, 0.0f};
  Targets tg;
  if (ok) {
    tg = targets_of(k);
    v = src[(int64_t)rr * g.ostride + p];
    if (conj) v.y = -v.y;
  }
#pragma unroll 1
  for (int q = 0; q < 8; ++q) {
    const int b = ok ? target_brick(tg, q, bg) : -1;
    const int slot = wave_claim<true>(b, cursor);
    if (b >= 0 && slot >= 0 && slot < cap) {
      const int cz = (q >> 2) & 1, cy = (q >> 1) & 1, cx = q & 1;
      // corner offset relative to the brick origin, in [-1, B-1]
      const int oz = k.z0 - (tg.blo[0] + cz) * kBZ, oy = k.y0 - (tg.blo[1] + cy) * kBY,
                ox = k.x0 - (tg.blo[2] + cx) * kBX;
      BinRec rc;
      rc.off = (unsigned)(oz + 1) | ((unsigned)(oy + 1) << 8) | ((unsigned)(ox + 1) << 16);
      rc.t[0] = k.tz;
      rc.t[1] = k.ty;
      rc.t[2] = k.tx;
      rc.v[0] = v.x;
      rc.v[1] = v.y;
      recs[slot] = rc;
    }
  }
}

// Exclusive prefix of the per-brick counts into offset and cursor (one block:
// each thread sums a contiguous run, a scan over the runs in LDS, then each
// thread writes its run's offsets).
__global__ __launch_bounds__(1024) void bin_scan_kernel(const int *__restrict__ count, int n,
                                                       int *__restrict__ offset, int *__restrict__ cursor) {
  __shared__ int s[1024];
  const int per = (n + 1023) / 1024;
  const int b0 = (int)threadIdx.x * per, b1 = min(b0 + per, n);
  int sum = 0;
  for (int b = b0; b < b1; ++b) sum += count[b];
  s[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int v = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += v;
    __syncthreads();
  }
  int run = s[threadIdx.x] - sum;
  for (int b = b0; b < b1; ++b) {
    offset[b] = run;
    cursor[b] = run;
    run += count[b];
  }
}

template <bool INSERT>
__global__ __launch_bounds__(kThreads) void bin_brick_kernel(SliceGeom g, BinGrid bg, const int *__restrict__ count,
                                                            const int *__restrict__ offset,
                                                            const BinRec *__restrict__ recs, f2v *__restrict__ out,
                                                            float *__restrict__ wout, f2v *__restrict__ plane,
                                                            float *__restrict__ wplane, int accumulate) {
  constexpr int NA = INSERT ? 3 : 2;
  __shared__ u64 tile[NA * kNV];
  __shared__ float s_max[kThreads / 64];
  const int b = blockIdx.x;
  const int bz = b / (bg.nby * bg.nbx), brem = b - bz * bg.nby * bg.nbx;
  const int by = brem / bg.nbx, bx = brem - by * bg.nbx;
  const int lo0 = bz * kBZ, lo1 = by * kBY, lo2 = bx * kBX;
  const int ext0 = min(kBZ, g.vD - lo0), ext1 = min(kBY, g.vH - lo1), ext2 = min(kBX, g.vW - lo2);
  const int cnt = count[b], off = offset[b];
  const int tid = threadIdx.x;
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  if (cnt == 0 && accumulate) {  // nothing to add; only the mirror plane needs its zeros
    if (INSERT && lo2 == 0)
      for (int v = tid; v < ext0 * ext1; v += kThreads) {
        const int64_t pe = (int64_t)(lo0 + v / ext1) * g.vH + lo1 + v % ext1;
        plane[pe] = f2v{0.0f, 0.0f};
        wplane[pe] = 0.0f;
      }
    return;
  }
  // the exact maximum |value| of the brick's records
  float m = 0.0f;
  for (int t = tid; t < cnt; t += kThreads) {
    const BinRec &rc = recs[off + t];
    m = fmaxf(m, fmaxf(fabsf(rc.v[0]), fabsf(rc.v[1])));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((tid & 63) == 0) s_max[tid >> 6] = m;
  for (int v = tid; v < NA * kNV; v += kThreads) tile[v] = 0ull;
  __syncthreads();
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) m = fmaxf(m, s_max[w]);
  // scale 2^(62 - ceil(log2 cnt) - e) with max < 2^e: a voxel sums at most
  // cnt contributions, each below 2^(62 - ceil(log2 cnt)) after scaling
  int e = 0;
  if (m > 0.0f && m < 3.0e38f) frexpf(m, &e);
  e = max(e, -60);
  const float vscale = ldexpf(1.0f, 62 - ceil_log2(cnt) - e);
  for (int t = tid; t < cnt; t += kThreads) {
    const BinRec rc = recs[off + t];
    const int oz = (int)(rc.off & 255u) - 1, oy = (int)((rc.off >> 8) & 255u) - 1,
              ox = (int)((rc.off >> 16) & 255u) - 1;
    const float vx = rc.v[0] * vscale, vy = rc.v[1] * vscale;
    const float wz[2] = {1.0f - rc.t[0], rc.t[0]}, wy[2] = {1.0f - rc.t[1], rc.t[1]},
                wx[2] = {1.0f - rc.t[2], rc.t[2]};
#pragma unroll
    for (int dz = 0; dz < 2; ++dz) {
      const int lz = oz + dz;
      if ((unsigned)lz >= (unsigned)kBZ) continue;
#pragma unroll
      for (int dy = 0; dy < 2; ++dy) {
        const int ly = oy + dy;
        if ((unsigned)ly >= (unsigned)kBY) continue;
        const float wzy = wz[dz] * wy[dy];
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int lx = ox + dx;
          if ((unsigned)lx >= (unsigned)kBX) continue;
          const float w = wzy * wx[dx];
          const int o = (lz * kBY + ly) * kBX + lx;
          atomicAdd(&tile[o], (u64)to_fixed64(__fmul_rn(w, vx)));
          atomicAdd(&tile[kNV + o], (u64)to_fixed64(__fmul_rn(w, vy)));
          if (INSERT && wout) atomicAdd(&tile[2 * kNV + o], (u64)(long long)to_fixed32(w, kWScale));
        }
      }
    }
  }
  __syncthreads();
  const float inv_v = 1.0f / vscale;
  for (int v = tid; v < kNV; v += kThreads) {
    const int lz = v / (kBY * kBX), ly = (v / kBX) % kBY, lx = v % kBX;
    if (lz >= ext0 || ly >= ext1 || lx >= ext2) continue;
    const int gz = lo0 + lz, gy = lo1 + ly, gx = lo2 + lx;
    const f2v val = {(float)(long long)tile[v] * inv_v, (float)(long long)tile[kNV + v] * inv_v};
    if (INSERT) {
      const float w = (float)(long long)tile[2 * kNV + v] * (1.0f / kWScale);
      if (gx == 0) {  // the x = 0 plane goes through the Friedel mirror pass
        plane[(int64_t)gz * g.vH + gy] = val;
        wplane[(int64_t)gz * g.vH + gy] = w;
        continue;
      }
      const int64_t ei = gz * sz + gy * sy + gx;
      if (wout) wout[ei] = accumulate ? wout[ei] + w : w;
      out[ei] = accumulate ? out[ei] + val : val;
    } else {
      const int64_t ei = gz * sz + gy * sy + gx;
      out[ei] = accumulate ? out[ei] + val : val;
    }
  }
}

__global__ __launch_bounds__(kThreads) void bin_zero_kernel(int *__restrict__ p, int n) {
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) p[i] = 0;
}

BinGrid bin_grid(const SliceGeom &g) {
  return BinGrid{(g.vD + kBZ - 1) / kBZ, (g.vH + kBY - 1) / kBY, (g.vW + kBX - 1) / kBX};
}

int64_t record_cap(const SliceGeom &g, int64_t n_rot) { return 8 * n_rot * (int64_t)g.P; }

bool binned_fits(const SliceGeom &g, int64_t n_rot) {
  const BinGrid bg = bin_grid(g);
  const int64_t n_bricks = (int64_t)bg.nbz * bg.nby * bg.nbx;
  return n_rot <= 65535 && n_bricks < ((int64_t)1 << 30) && record_cap(g, n_rot) < ((int64_t)1 << 31) &&
         record_cap(g, n_rot) * (int64_t)sizeof(BinRec) <= kMaxRecordBytes;
}

}  // namespace

int set_scatter_path(int path) {
  if (path != TT_SCATTER_AUTO && path != TT_SCATTER_BRICK && path != TT_SCATTER_BINNED)
    return fail_invalid("scatter path must be TT_SCATTER_AUTO, TT_SCATTER_BRICK or TT_SCATTER_BINNED");
  g_scatter_path = path;
  return TT_OK;
}

bool binned_selected(const SliceGeom &g, int64_t n_rot) {
  if (g_scatter_path == TT_SCATTER_BRICK || !binned_fits(g, n_rot)) return false;
  if (g_scatter_path == TT_SCATTER_BINNED) return true;
  const BinGrid bg = bin_grid(g);
  const double n_bricks = (double)bg.nbz * bg.nby * bg.nbx;
  return (double)n_rot * (double)g.P <= kAutoPixelsPerBrick * n_bricks;
}

int binned_scatter(bool insert, const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                   float *wout, int accumulate, hipStream_t s) {
  if (!binned_fits(g, n_rot)) return fail_invalid("binned scatter: too many slice pixels for the record buffer");
  const BinGrid bg = bin_grid(g);
  const int n_bricks = bg.nbz * bg.nby * bg.nbx;
  const int64_t cap = record_cap(g, n_rot);
  const size_t np = (size_t)g.vD * g.vH;
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_count = 0, o_offset = up(o_count + (size_t)n_bricks * 4), o_cursor = up(o_offset + (size_t)n_bricks * 4);
  const size_t o_plane = up(o_cursor + (size_t)n_bricks * 4);
  const size_t o_wplane = up(o_plane + (insert ? np * sizeof(f2v) : 0));
  const size_t o_recs = up(o_wplane + (insert ? np * sizeof(float) : 0));
  const size_t bytes = o_recs + (size_t)std::max<int64_t>(cap, 1) * sizeof(BinRec);
  void *base = nullptr;
  if (hipError_t e = hipMallocAsync(&base, bytes, s)) {
    set_error(std::string("binned scatter: hipMallocAsync: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  char *c = (char *)base;
  int *count = (int *)(c + o_count), *offset = (int *)(c + o_offset), *cursor = (int *)(c + o_cursor);
  f2v *plane = insert ? (f2v *)(c + o_plane) : nullptr;
  float *wplane = insert ? (float *)(c + o_wplane) : nullptr;
  BinRec *recs = (BinRec *)(c + o_recs);
  int st = TT_OK;
  hipLaunchKernelGGL(bin_zero_kernel, dim3((unsigned)std::min((n_bricks + kThreads - 1) / kThreads, 4096)),
                     dim3(kThreads), 0, s, count, n_bricks);
  st = check_launch("bin_zero_kernel");
  const dim3 pgrid((unsigned)((g.P + kThreads - 1) / kThreads), (unsigned)n_rot);
  if (!st && n_rot > 0) {
    if (insert)
      hipLaunchKernelGGL(bin_count_kernel<true>, pgrid, dim3(kThreads), 0, s, g, rot, bg, count);
    else
      hipLaunchKernelGGL(bin_count_kernel<false>, pgrid, dim3(kThreads), 0, s, g, rot, bg, count);
    st = check_launch("bin_count_kernel");
  }
  if (!st) {
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, s, (const int *)count, n_bricks, offset, cursor);
    st = check_launch("bin_scan_kernel");
  }
  if (!st && n_rot > 0) {
    if (insert)
      hipLaunchKernelGGL(bin_fill_kernel<true>, pgrid, dim3(kThreads), 0, s, g, src, rot, bg, cursor, recs, cap);
    else
      hipLaunchKernelGGL(bin_fill_kernel<false>, pgrid, dim3(kThreads), 0, s, g, src, rot, bg, cursor, recs, cap);
    st = check_launch("bin_fill_kernel");
  }
  if (!st) {
    if (insert)
      hipLaunchKernelGGL(bin_brick_kernel<true>, dim3((unsigned)n_bricks), dim3(kThreads), 0, s, g, bg,
                         (const int *)count, (const int *)offset, (const BinRec *)recs, out, wout, plane, wplane,
                         accumulate);
    else
      hipLaunchKernelGGL(bin_brick_kernel<false>, dim3((unsigned)n_bricks), dim3(kThreads), 0, s, g, bg,
                         (const int *)count, (const int *)offset, (const BinRec *)recs, out, wout, plane, wplane,
                         accumulate);
    st = check_launch("bin_brick_kernel");
  }
  if (!st && insert) st = launch_mirror(plane, wplane, g, out, wout, accumulate, s);
  const hipError_t fe = hipFreeAsync(base, s);
  if (!st && fe != hipSuccess) {
    set_error(std::string("binned scatter: hipFreeAsync: ") + hipGetErrorString(fe));
    st = TT_ERR_HIP;
  }
  return st;
}

}  // namespace tt
