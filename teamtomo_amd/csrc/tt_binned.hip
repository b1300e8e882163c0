// Binned scatter: the extraction adjoint and the insertion when the slice
// pixels are few per volume brick, as for a tilt series inserted into a large
// volume (61 slices of 512^2 into 512^3: ~196 pixels per 4x8x64 brick).
//
// The brick engine discovers each brick's slices itself (plane tests, FP64
// item frames, row records, chunk steps), a chain of dependent steps that costs
// more than the few pixels it then scatters; its upper-shell halos also make
// a second pass read and rewrite much of the volume.  Here the
// pixels come to the bricks instead:
//   1. bin_count_kernel: one lane per pixel evaluates the canonical FP64
//      coordinate (the pixel-atomic kernels' code, tt_common.h) and counts it
//      in every brick that owns one of its 8 trilinear corners with a
//      nonzero weight (1, 2, 4 or 8 bricks); a block takes a 32 x 32 pixel
//      patch and counts in an LDS table first, one global atomic per brick;
//   2. bin_scan_kernel: exclusive prefix of the counts;
//   3. bin_fill_kernel: the same evaluation again; the block reserves a run
//      of each bin it reaches and writes each pixel as a 24-byte record
//      (corner offset, fractions, value) into its bricks' bins, and raises
//      the bricks' maxima |value|;
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
#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_engine.h"
#include "tt_fixed.h"

namespace tt {

namespace {

// Voxels a brick owns per axis: x rows of 64 voxels (512 bytes of the output
// row) keep most of the flush's cache lines whole; the rfft rows of 2^k + 1
// voxels leave every brick row misaligned, so shorter rows write two partial
// lines per row (61 tilts of 512^2 into 512^3, ms: 8x16x16 0.80, 4x8x32
// 0.74, 4x8x64 0.70, 2x8x128 0.82).
constexpr int kBZ = 4, kBY = 8, kBX = 64;
// The last brick along x also owns the volume's final x voxel when the
// other bricks leave one over (rfft rows of 2^k + 1 voxels): its tile rows
// hold kBX + 1 voxels, and no block serves a one-voxel column.
constexpr int kBXT = kBX + 1;
constexpr int kNV = kBZ * kBY * kBXT;
constexpr int kThreads = 256;       // pixel passes
// brick pass: 512 threads, three blocks of 8 waves per CU with ~50 KB of LDS
// each (61 tilts of 512^2 into 512^3: 256 threads 0.669 ms, 512 0.616 ms,
// 1024 0.659 ms)
constexpr int kBrickThreads = 512;
// A pixel lands in at most 8 bricks: the record buffer holds 8 per pixel (a
// call writes about 1.25 per pixel), and the path is used only while that
// buffer stays below this size; if the allocation fails, the call falls back
// to the brick engine.
constexpr int64_t kMaxRecordBytes = (int64_t)8 << 30;
// Auto selection: at most this many pixels (slices x pixels per slice) per
// brick of the volume on average.  Measured (insertion / adjoint, ms, brick
// engine against binned): 61 tilts of 512^2 into 512^3 (196 per brick) 2.30 /
// 1.46 against 0.72 / 0.55; 61 of 256^2 into 256^3 (328) 0.39 / 0.28 against
// 0.19 / 0.17; 64 random slices into 256^3 (344) 0.39 / 0.25 against 0.22 /
// 0.16; 256 into 256^3 (1376) 0.48 / 0.31 against 0.55 / 0.49; 256 into 128^3
// (2080) 0.15 / 0.16 against 0.31 / 0.27 (the bins near DC then take most of
// the records).  The insertions of the 32 volumes of the batch workload (64^3
// to 192^3, 16 to 2048 slices) cross over between 700 and 950 per brick.
constexpr double kAutoPixelsPerBrick = 800.0;

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

// The record buffers come from a memory pool of this library per device that
// keeps up to kPoolKeep bytes mapped between calls: buffers of hundreds of
// MB allocated and freed by every call would otherwise be mapped again each
// time (a loop of small binned calls measured 3 ms slower than with the brick
// engine until they were pooled).
constexpr uint64_t kPoolKeep = (uint64_t)4 << 30;
constexpr int kMaxPoolDevices = 64;
std::mutex g_pool_mutex;
hipMemPool_t g_pools[kMaxPoolDevices] = {};

// The pool of the current device, created at its first use; nullptr if the
// device has no pools (then the default pool is used).
hipMemPool_t bin_pool() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxPoolDevices) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  if (!g_pools[dev]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    if (hipMemPoolCreate(&pool, &props) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    uint64_t keep = kPoolKeep;
    if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) != hipSuccess) (void)hipGetLastError();
    g_pools[dev] = pool;
  }
  return g_pools[dev];
}

// A slice pixel's cell, as the pixel-atomic scatter evaluates it; false for a
// pixel the scatter skips (the insertion's x = 0, y < 0 half row, outside the
// disc or the volume).
template <bool INSERT>
__device__ __forceinline__ bool bin_pixel(const SliceGeom &g, const Rot &r, int i, int j, Cell &k, bool &conj) {
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

__device__ __forceinline__ Targets targets_of(const Cell &k, const BinGrid &bg) {
  Targets tg;
  const int c[3] = {k.z0, k.y0, k.x0};
  const float t[3] = {k.tz, k.ty, k.tx};
  const int B[3] = {kBZ, kBY, kBX};
  tg.lo_ok = tg.hi_ok = 0u;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    int lo = c[d] / B[d], hi = (c[d] + 1) / B[d];
    if (d == 2) {  // the last x brick owns the final voxel
      lo = min(lo, bg.nbx - 1);
      hi = min(hi, bg.nbx - 1);
    }
    tg.blo[d] = lo;
    if (t[d] < 1.0f || hi == lo) tg.lo_ok |= 1u << d;
    if (hi != lo && t[d] > 0.0f) tg.hi_ok |= 1u << d;
  }
  return tg;
}

// The targets q (bit 2: upper on z, bit 1: on y, bit 0: on x) as a bit mask,
// so that the pixel loops visit only the bricks a cell reaches (one for
// about two thirds of the pixels) rather than test all eight.
__device__ __forceinline__ unsigned target_mask(const Targets &tg) {
  const unsigned z = ((tg.lo_ok & 1u) ? 0x0fu : 0u) | ((tg.hi_ok & 1u) ? 0xf0u : 0u);
  const unsigned y = ((tg.lo_ok & 2u) ? 0x33u : 0u) | ((tg.hi_ok & 2u) ? 0xccu : 0u);
  const unsigned x = ((tg.lo_ok & 4u) ? 0x55u : 0u) | ((tg.hi_ok & 4u) ? 0xaau : 0u);
  return z & y & x;
}

// Brick of target q (a set bit of target_mask).
__device__ __forceinline__ int target_brick(const Targets &tg, int q, const BinGrid &bg) {
  const int cz = (q >> 2) & 1, cy = (q >> 1) & 1, cx = q & 1;
  return ((tg.blo[0] + cz) * bg.nby + (tg.blo[1] + cy)) * bg.nbx + (tg.blo[2] + cx);
}

// Pixel passes: a block takes a kPR x kPC patch of one slice, whose pixels
// fall in a compact region of the volume.  Its (pixel, brick) pairs are
// counted per brick in an LDS hash table, and the block then makes one global
// atomic per brick it reaches instead of one per pair (pairs beyond the
// table's capacity go straight to the global counters).
constexpr int kPR = 32, kPC = 32, kPerThread = kPR * kPC / kThreads;
constexpr int kTableBits = 9, kTable = 1 << kTableBits;

__device__ __forceinline__ int table_hash(int key) {
  return (int)(((unsigned)key * 2654435761u) >> (32 - kTableBits));
}

// The entry of key, inserted if absent; -1 if the table is full.
__device__ __forceinline__ int table_insert(int *__restrict__ keys, int key) {
  int h = table_hash(key);
  for (int probe = 0; probe < kTable; ++probe) {
    const int old = atomicCAS(&keys[h], -1, key);
    if (old == -1 || old == key) return h;
    h = (h + 1) & (kTable - 1);
  }
  return -1;
}

// The entry of key, or -1 if it is not in the table (linear probing, no
// deletions: the probe sequence of a present key has no empty entry before it).
__device__ __forceinline__ int table_find(const int *__restrict__ keys, int key) {
  int h = table_hash(key);
  for (int probe = 0; probe < kTable; ++probe) {
    const int k = keys[h];
    if (k == key) return h;
    if (k == -1) return -1;
    h = (h + 1) & (kTable - 1);
  }
  return -1;
}

// Pixel u of this thread in the block's patch: (i, j), false past the slice.
__device__ __forceinline__ bool patch_pixel(const SliceGeom &g, int u, int &i, int &j) {
  const int npc = (g.Wr + kPC - 1) / kPC;
  const int pi = (int)blockIdx.x / npc, pj = (int)blockIdx.x - pi * npc;
  const int l = (int)threadIdx.x + u * kThreads;
  i = pi * kPR + l / kPC;
  j = pj * kPC + l % kPC;
  return i < g.H && j < g.Wr;
}

template <bool INSERT>
__global__ __launch_bounds__(kThreads) void bin_count_kernel(SliceGeom g, const float *__restrict__ rot,
                                                            BinGrid bg, int *__restrict__ count) {
  __shared__ int s_key[kTable], s_cnt[kTable];
  for (int e = threadIdx.x; e < kTable; e += kThreads) {
    s_key[e] = -1;
    s_cnt[e] = 0;
  }
  __syncthreads();
  const Rot r = load_rot(rot + (int64_t)blockIdx.y * 9, g);
#pragma unroll 1
  for (int u = 0; u < kPerThread; ++u) {
    int i, j;
    Cell k;
    bool conj;
    if (!patch_pixel(g, u, i, j) || !bin_pixel<INSERT>(g, r, i, j, k, conj)) continue;
    const Targets tg = targets_of(k, bg);
#pragma unroll 1
    for (unsigned tm = target_mask(tg); tm; tm &= tm - 1u) {
      const int q = __builtin_ctz(tm);
      const int b = target_brick(tg, q, bg);
      const int h = table_insert(s_key, b);
      if (h >= 0)
        atomicAdd(&s_cnt[h], 1);
      else
        atomicAdd(&count[b], 1);
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kTable; e += kThreads)
    if (s_key[e] >= 0) atomicAdd(&count[s_key[e]], s_cnt[e]);
}

// Writes the records: the block reserves, per brick it reaches, a run of
// slots in the brick's bin with one global atomic (and raises the brick's
// maximum |value|), then its pairs take consecutive slots of their runs.
template <bool INSERT>
__global__ __launch_bounds__(kThreads) void bin_fill_kernel(SliceGeom g, const f2v *__restrict__ src,
                                                           const float *__restrict__ rot, BinGrid bg,
                                                           int *__restrict__ cursor, unsigned *__restrict__ bmax,
                                                           BinRec *__restrict__ recs, int64_t cap) {
  __shared__ int s_key[kTable], s_cnt[kTable], s_base[kTable];
  __shared__ unsigned s_mx[kTable];
  for (int e = threadIdx.x; e < kTable; e += kThreads) {
    s_key[e] = -1;
    s_cnt[e] = 0;
    s_mx[e] = 0u;
  }
  __syncthreads();
  const int rr = blockIdx.y;
  const Rot r = load_rot(rot + (int64_t)rr * 9, g);
  Cell cell[kPerThread];
  f2v val[kPerThread];
  unsigned okm = 0u;
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    int i, j;
    bool conj = false;
    if (patch_pixel(g, u, i, j) && bin_pixel<INSERT>(g, r, i, j, cell[u], conj)) {
      okm |= 1u << u;
      f2v v = src[(int64_t)rr * g.ostride + (int64_t)i * g.Wr + j];
      if (conj) v.y = -v.y;
      val[u] = v;
    }
  }
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    if (!((okm >> u) & 1u)) continue;
    const Targets tg = targets_of(cell[u], bg);
    const unsigned vb = __float_as_uint(fmaxf(fabsf(val[u].x), fabsf(val[u].y)));
#pragma unroll 1
    for (unsigned tm = target_mask(tg); tm; tm &= tm - 1u) {
      const int q = __builtin_ctz(tm);
      const int b = target_brick(tg, q, bg);
      const int h = table_insert(s_key, b);
      if (h >= 0) {
        atomicAdd(&s_cnt[h], 1);
        atomicMax(&s_mx[h], vb);
      }
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kTable; e += kThreads) {
    const int key = s_key[e];
    if (key < 0) continue;
    s_base[e] = atomicAdd(&cursor[key], s_cnt[e]);
    atomicMax(&bmax[key], s_mx[e]);
    s_cnt[e] = 0;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kPerThread; ++u) {
    if (!((okm >> u) & 1u)) continue;
    const Cell &k = cell[u];
    const Targets tg = targets_of(k, bg);
#pragma unroll 1
    for (unsigned tm = target_mask(tg); tm; tm &= tm - 1u) {
      const int q = __builtin_ctz(tm);
      const int b = target_brick(tg, q, bg);
      const int h = table_find(s_key, b);
      int slot;
      if (h >= 0) {
        slot = s_base[h] + atomicAdd(&s_cnt[h], 1);
      } else {  // the table was full when this pair was counted
        slot = atomicAdd(&cursor[b], 1);
        atomicMax(&bmax[b], __float_as_uint(fmaxf(fabsf(val[u].x), fabsf(val[u].y))));
      }
      if (slot < 0 || slot >= cap) continue;
      const int cz = (q >> 2) & 1, cy = (q >> 1) & 1, cx = q & 1;
      // corner offset relative to the brick origin, in [-1, B-1]
      const int oz = k.z0 - (tg.blo[0] + cz) * kBZ, oy = k.y0 - (tg.blo[1] + cy) * kBY,
                ox = k.x0 - (tg.blo[2] + cx) * kBX;
      BinRec rc;
      rc.off = (unsigned)(oz + 1) | ((unsigned)(oy + 1) << 8) | ((unsigned)(ox + 1) << 16);
      rc.t[0] = k.tz;
      rc.t[1] = k.ty;
      rc.t[2] = k.tx;
      rc.v[0] = val[u].x;
      rc.v[1] = val[u].y;
      recs[slot] = rc;
    }
  }
}

// Exclusive prefix of the per-brick counts into offset and cursor, one block
// of 1024 threads over chunks of 8192 counts staged in LDS (coalesced loads
// and stores): each thread sums a run of 8, the runs are scanned with DPP
// lane moves inside each wave and the 16 wave totals, then the runs'
// prefixes are written back.
__global__ __launch_bounds__(1024) void bin_scan_kernel(const int *__restrict__ count, int n,
                                                       int *__restrict__ offset, int *__restrict__ cursor) {
  constexpr int kPer = 8, kChunk = 1024 * kPer;
  __shared__ int s_v[kChunk];
  __shared__ int s_wave[16];
  const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
  int carry = 0;
  for (int c0 = 0; c0 < n; c0 += kChunk) {
    for (int e = tid; e < kChunk; e += 1024) s_v[e] = c0 + e < n ? count[c0 + e] : 0;
    __syncthreads();
    int run = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) run += s_v[tid * kPer + q];
    const int incl = wave_incl_scan(run);
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    int wbase = 0, total = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int t = s_wave[k];
      wbase += k < wave ? t : 0;
      total += t;
    }
    int x = carry + wbase + incl - run;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int c = s_v[tid * kPer + q];
      s_v[tid * kPer + q] = x;
      x += c;
    }
    __syncthreads();
    for (int e = tid; e < kChunk; e += 1024)
      if (c0 + e < n) {
        offset[c0 + e] = s_v[e];
        cursor[c0 + e] = s_v[e];
      }
    carry += total;
    __syncthreads();  // the next chunk rewrites s_v and the wave totals
  }
}

template <bool INSERT>
__global__ __launch_bounds__(kBrickThreads) void bin_brick_kernel(SliceGeom g, BinGrid bg,
                                                                 const int *__restrict__ count,
                                                                 const int *__restrict__ offset,
                                                                 const unsigned *__restrict__ bmax,
                                                                 const BinRec *__restrict__ recs,
                                                                 f2v *__restrict__ out, float *__restrict__ wout,
                                                                 f2v *__restrict__ plane, float *__restrict__ wplane,
                                                                 int accumulate) {
  constexpr int NA = INSERT ? 3 : 2;
  __shared__ u64 tile[NA * kNV];
  // (grouping the x bricks of a brick row on one XCD, so that the partial
  // cache lines they share meet in one L2, measured 5 % slower)
  const int b = blockIdx.x;
  const int row = b / bg.nbx, bx = b - row * bg.nbx;
  const int bz = row / bg.nby, by = row - bz * bg.nby;
  const int lo0 = bz * kBZ, lo1 = by * kBY, lo2 = bx * kBX;
  const int ext0 = min(kBZ, g.vD - lo0), ext1 = min(kBY, g.vH - lo1);
  const int ext2 = bx == bg.nbx - 1 ? g.vW - lo2 : kBX;
  const int cnt = count[b], off = offset[b];
  const float m = __uint_as_float(bmax[b]);  // the exact maximum |value| of the brick's records
  const int tid = threadIdx.x;
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  if (cnt == 0 && accumulate) {  // nothing to add; only the mirror plane needs its zeros
    if (INSERT && lo2 == 0)
      for (int v = tid; v < ext0 * ext1; v += kBrickThreads) {
        const int64_t pe = (int64_t)(lo0 + v / ext1) * g.vH + lo1 + v % ext1;
        plane[pe] = f2v{0.0f, 0.0f};
        wplane[pe] = 0.0f;
      }
    return;
  }
  for (int v = tid; v < NA * kNV; v += kBrickThreads) tile[v] = 0ull;
  // scale 2^(62 - ceil(log2 cnt) - e) with max < 2^e: a voxel sums at most
  // cnt contributions, each below 2^(62 - ceil(log2 cnt)) after scaling
  int e = 0;
  if (m > 0.0f && m < 3.0e38f) frexpf(m, &e);
  e = max(e, -60);
  const float vscale = ldexpf(1.0f, 62 - ceil_log2(cnt) - e);
  __syncthreads();
  for (int t = tid; t < cnt; t += kBrickThreads) {
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
          if ((unsigned)lx >= (unsigned)ext2) continue;
          const float w = wzy * wx[dx];
          const int o = (lz * kBY + ly) * kBXT + lx;
          atomicAdd(&tile[o], (u64)to_fixed64(__fmul_rn(w, vx)));
          atomicAdd(&tile[kNV + o], (u64)to_fixed64(__fmul_rn(w, vy)));
          if (INSERT && wout) atomicAdd(&tile[2 * kNV + o], (u64)(long long)to_fixed32(w, kWScale));
        }
      }
    }
  }
  __syncthreads();
  const float inv_v = 1.0f / vscale;
  // a wave writes one 64-voxel x row; the last brick's extra column follows
  for (int u = tid; u < kBZ * kBY * (kBX + 1); u += kBrickThreads) {
    const bool extra = u >= kBZ * kBY * kBX;
    const int r = extra ? u - kBZ * kBY * kBX : u / kBX;  // x row lz kBY + ly
    const int lz = r / kBY, ly = r % kBY, lx = extra ? kBX : u % kBX;
    if (lz >= ext0 || ly >= ext1 || lx >= ext2) continue;
    const int v = r * kBXT + lx;
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
  // x: bricks of kBX voxels, the last one kBX + 1 wide when that covers the row
  return BinGrid{(g.vD + kBZ - 1) / kBZ, (g.vH + kBY - 1) / kBY, std::max(1, (g.vW - 1 + kBX - 1) / kBX)};
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
  // pixels per 4 x 8 x 64 block of the volume (the measure the threshold was
  // measured in, whether or not the last x brick is wider)
  const BinGrid bg = bin_grid(g);
  const double n_blocks = (double)bg.nbz * bg.nby * (double)((g.vW + kBX - 1) / kBX);
  return (double)n_rot * (double)g.P <= kAutoPixelsPerBrick * n_blocks;
}

int binned_scatter(bool insert, const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                   float *wout, int accumulate, hipStream_t s) {
  if (!binned_fits(g, n_rot)) return fail_invalid("binned scatter: too many slice pixels for the record buffer");
  const BinGrid bg = bin_grid(g);
  const int n_bricks = bg.nbz * bg.nby * bg.nbx;
  const int64_t cap = record_cap(g, n_rot);
  const size_t np = (size_t)g.vD * g.vH;
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  // count and bmax adjacent: one zero pass clears both
  const size_t o_count = 0, o_bmax = o_count + (size_t)n_bricks * 4, o_offset = up(o_bmax + (size_t)n_bricks * 4),
               o_cursor = up(o_offset + (size_t)n_bricks * 4);
  const size_t o_plane = up(o_cursor + (size_t)n_bricks * 4);
  const size_t o_wplane = up(o_plane + (insert ? np * sizeof(f2v) : 0));
  const size_t o_recs = up(o_wplane + (insert ? np * sizeof(float) : 0));
  // sizes rounded up to a power of two, so that calls of similar sizes reuse
  // the pool's blocks
  size_t bytes = o_recs + (size_t)std::max<int64_t>(cap, 1) * sizeof(BinRec);
  bytes = (size_t)1 << ceil_log2((int64_t)bytes);
  void *base = nullptr;
  hipMemPool_t pool = bin_pool();
  if (hipError_t e = pool ? hipMallocFromPoolAsync(&base, bytes, pool, s) : hipMallocAsync(&base, bytes, s)) {
    if (e == hipErrorOutOfMemory) {  // the caller runs the brick engine instead
      (void)hipGetLastError();
      return kBinnedNoMemory;
    }
    set_error(std::string("binned scatter: hipMallocAsync: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  char *c = (char *)base;
  int *count = (int *)(c + o_count), *offset = (int *)(c + o_offset), *cursor = (int *)(c + o_cursor);
  unsigned *bmax = (unsigned *)(c + o_bmax);
  f2v *plane = insert ? (f2v *)(c + o_plane) : nullptr;
  float *wplane = insert ? (float *)(c + o_wplane) : nullptr;
  BinRec *recs = (BinRec *)(c + o_recs);
  int st = TT_OK;
  hipLaunchKernelGGL(bin_zero_kernel, dim3((unsigned)std::min((2 * n_bricks + kThreads - 1) / kThreads, 4096)),
                     dim3(kThreads), 0, s, count, 2 * n_bricks);
  st = check_launch("bin_zero_kernel");
  const dim3 pgrid((unsigned)(((g.H + kPR - 1) / kPR) * ((g.Wr + kPC - 1) / kPC)), (unsigned)n_rot);
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
      hipLaunchKernelGGL(bin_fill_kernel<true>, pgrid, dim3(kThreads), 0, s, g, src, rot, bg, cursor, bmax, recs,
                         cap);
    else
      hipLaunchKernelGGL(bin_fill_kernel<false>, pgrid, dim3(kThreads), 0, s, g, src, rot, bg, cursor, bmax, recs,
                         cap);
    st = check_launch("bin_fill_kernel");
  }
  if (!st) {
    const unsigned bgrid = (unsigned)n_bricks;
    if (insert)
      hipLaunchKernelGGL(bin_brick_kernel<true>, dim3(bgrid), dim3(kBrickThreads), 0, s, g, bg,
                         (const int *)count, (const int *)offset, (const unsigned *)bmax, (const BinRec *)recs, out,
                         wout, plane, wplane, accumulate);
    else
      hipLaunchKernelGGL(bin_brick_kernel<false>, dim3(bgrid), dim3(kBrickThreads), 0, s, g, bg,
                         (const int *)count, (const int *)offset, (const unsigned *)bmax, (const BinRec *)recs, out,
                         wout, plane, wplane, accumulate);
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
