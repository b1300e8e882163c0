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
//           zeroed by row blocks spread over the same launch, which write
//           exactly the pixels no brick writes.
//  scatter: the brick accumulates in LDS in fixed point with integer atomics
//           (8 corners per pixel, the upper corners land in a halo shell),
//           then stores its own voxels and passes the shell to the
//           neighbouring bricks through a halo buffer.
//
// Inside a workgroup the waves are independent: each takes a contiguous share
// of the unit's rotations and runs
//  (1) discovery: one lane per rotation tests the slice plane against the
//      brick's box; per crossing rotation and Friedel half, an item holds the
//      range of candidate slice rows and an FP32 frame referred to the tile
//      centre;
//  (2) rows: one lane per row evaluates the row's FP32 frame coordinate and
//      the interval of columns whose coordinates can lie in the box, and
//      writes the row's record;
//  (3) chunks: the rows' columns are cut into chunks of CH pixels, handed out
//      64 at a time, consecutive lanes on consecutive chunks of a row (a lane
//      finds its row by a running maximum over marks that the rows write at
//      their first chunk).  A pixel keeps its FP32 cell if no coordinate lies
//      near an integer, else evaluates its canonical FP64 cell, and is taken
//      only if that cell is in the brick.
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
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_engine.h"
#include "tt_fixed.h"

namespace tt {

namespace {

enum Mode { kGather = 0, kAdjoint = 1, kInsert = 2 };

constexpr double kBoxMargin = 1e-6;  // widening of the candidate box (voxels)

// Per-rotation plane data, shared by all bricks.  The slice plane of pixel
// (m, j), m = i - H//2, is x = beta + m A + j B with A = U / H, B = V (array
// units); (As, Bs) is the dual basis of (A, B) in the plane, so that a point x
// of the plane has m = (x - beta) . As and j = (x - beta) . Bs.
struct PlaneRec {
  double U[3], V[3];    // canonical constants (tt_common.h)
  double As[3], Bs[3];  // dual basis (0 for a degenerate plane)
  int tie;              // extraction conj decision on the x' = 0 plane
  int full;             // degenerate plane: every row and column is a candidate
};

// Plane test data in FP32: n = the unit normal of span(A, B), n.w < 0
// marking a degenerate plane (every brick is a candidate); b = the column
// dual vector B* (the FP64 one of PlaneRec rounded), which tells a brick which
// Friedel halves of the slice can reach it (the sign of j = (x - beta) . s B*
// over its box).
struct PlaneTest {
  float4 n, b;
};

// Bits of a bound on the pixels of one slice that contribute to one voxel:
// their coordinates lie in the open cube of side 2 around it, a plane section
// of diameter at most D = 2 sqrt(3); the pixel lattice has column step |B|
// and row spacing h = |A x B| / |B| in the plane, so at most
// (D / h + 1)(D / |B| + 1) pixels fit, and never more than the slice's P.
__device__ __forceinline__ int slice_count_bits(const SliceGeom &g, double nn, double bb) {
  const double D = 3.4641016151377544;
  double cnt = (double)g.P;
  if (nn > 0.0 && bb > 0.0) {
    const double lb = sqrt(bb);
    cnt = fmin(cnt, (D * lb / nn + 1.0) * (D / lb + 1.0));
  }
  int e = 0;
  frexp(cnt, &e);  // cnt < 2^e
  return e;
}

__device__ __forceinline__ int plane_rec_one(const SliceGeom &g, const float *__restrict__ rot, int r,
                                             PlaneRec *__restrict__ out, PlaneTest *__restrict__ test) {
  const Rot rr = load_rot(rot + (int64_t)r * 9, g);
  PlaneRec p;
  double A[3], B[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    p.U[d] = rr.U[d];
    p.V[d] = rr.V[d];
    A[d] = rr.U[d] * g.inv_H;
    B[d] = rr.V[d];
  }
  const double aa = A[0] * A[0] + A[1] * A[1] + A[2] * A[2];
  const double bb = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
  const double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
  const double det = aa * bb - ab * ab;
  const bool full = !(det > 1e-12 * aa * bb) || !(aa > 0.0) || !(bb > 0.0);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    p.As[d] = full ? 0.0 : (bb * A[d] - ab * B[d]) / det;
    p.Bs[d] = full ? 0.0 : (aa * B[d] - ab * A[d]) / det;
  }
  p.tie = rr.tie ? 1 : 0;
  p.full = full ? 1 : 0;
  out[r] = p;
  double n0 = A[1] * B[2] - A[2] * B[1];
  double n1 = A[2] * B[0] - A[0] * B[2];
  double n2 = A[0] * B[1] - A[1] * B[0];
  const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  PlaneTest t;
  if (full || !(nn > 0.0)) {
    t.n = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
  } else {
    t.n = make_float4((float)(n0 / nn), (float)(n1 / nn), (float)(n2 / nn), 1.0f);
  }
  t.b = make_float4((float)p.Bs[0], (float)p.Bs[1], (float)p.Bs[2], 0.0f);
  test[r] = t;
  return slice_count_bits(g, nn, bb);
}

// The wave's largest per-slice count bits into *bits (if given).
__device__ __forceinline__ void max_count_bits(int b, unsigned *bits) {
  for (int off = 32; off > 0; off >>= 1) b = max(b, __shfl_xor(b, off, 64));
  if (bits && (threadIdx.x & 63) == 0 && b > 0) atomicMax(bits, (unsigned)b);
}

__global__ __launch_bounds__(256) void plane_rec_kernel(SliceGeom g, const float *__restrict__ rot,
                                                        int n_rot, PlaneRec *__restrict__ out,
                                                        PlaneTest *__restrict__ test, unsigned *bits) {
  const int r = blockIdx.x * 256 + threadIdx.x;
  const int b = r < n_rot ? plane_rec_one(g, rot, r, out, test) : 0;
  max_count_bits(b, bits);
}

// ---------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & 63); }

__device__ __forceinline__ float uniform_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}

// A wave-uniform double moved to scalar registers (VALU results live in vector
// registers even when every lane holds the same value).
__device__ __forceinline__ double uniform_d(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ __forceinline__ int rank_below(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
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
  // pixels per chunk: three for the adjoint, two otherwise (measured at
  // 256^3 x 4096: adjoint 0.991 ms with three against 1.007 ms with two;
  // three cost the insertion 2 % and the gather 5 %, where they spill 68
  // bytes per lane; four slowed both; one pixel per lane makes the scatter
  // 20 % slower)
  static constexpr int CH = (MODE == kAdjoint) ? 3 : 2;
  // tile: the brick's cells lo..lo+B-1 and their corners, B + 1 voxels per axis
  static constexpr int SX = BX + 1, SY = BY + 1, SZ = BZ + 1;
  static constexpr int NT = SX * SY * SZ;
  static constexpr int NACC = (MODE == kGather) ? 0 : ((MODE == kInsert) ? 3 : 2);
  // upper shell of a scatter tile (voxels l_z = BZ, or l_y = BY, or l_x = BX),
  // which belongs to the neighbouring bricks: Z face, then Y face, then X face
  static constexpr int NSH = SY * SX + BZ * SX + BZ * BY;
};

template <int BZ, int BY, int BX>
__device__ __forceinline__ int shell_index(int lz, int ly, int lx) {
  constexpr int SX = BX + 1, SY = BY + 1;
  if (lz == BZ) return ly * SX + lx;
  if (ly == BY) return SY * SX + lz * SX + lx;
  return SY * SX + BZ * SX + lz * BY + ly;
}


// A slice's item in a brick: a rotation and Friedel half whose rows can reach
// the brick, with its FP32 frame.  The coordinate of pixel (i, j) relative to
// the tile centre is st + (i - i_ref) sa + (j - j_ref) sv; st is the
// canonical FP64 coordinate of the reference pixel (the one nearest the tile
// centre) rounded to FP32, the steps are rounded from FP64.  Over the ~30
// rows and columns of a tile this stays within ~5e-6 voxel of the canonical
// value (about 1e-6 in practice); pixels within kNear of an integer
// coordinate re-evaluate the exact canonical FP64 coordinate instead.
struct __align__(16) ItemRec {
  float st[3];   // tile-local coordinate of the reference pixel
  float sa[3];   // step per slice row
  float sv[3];   // step per slice column
  int rows;      // first row i0 | row count << 16
  unsigned ref;  // reference row i_ref | reference column j_ref << 16
  unsigned rnt;  // rotation | tie << 30 | Friedel half (s = -1) << 31
};

// A row of a row pass, written once by the row's lane and read by each of its
// chunks: the row's FP32 frame coordinate at column jref (rowc, the same FMA
// the row's interval test evaluates), the column step, and where its
// candidate columns start.
struct __align__(8) RowRec {
  float rowc[3];  // frame coordinate of (i, jref)
  float sv[3];    // step per slice column
  float djl;      // jl - jref: the first candidate column relative to jref
  unsigned ijl;   // slice row i | first candidate column jl << 16
  int len;        // candidate columns from jl
  unsigned rnt;   // rotation | tie << 30 | Friedel half (s = -1) << 31
};

constexpr float kRowMargin = 1e-4f;    // widening of the row phase's tile box (voxels)

// Fixed-point scale of the scatter, one per brick: values are multiplied by
// 2^(30 - headroom - e), with 2^e > the largest |v| the brick can receive,
// so a value up to 2^headroom times that maximum still converts to a 32-bit
// integer; a larger (or non-finite) value raises *flag and the call is redone
// with the exact maximum of the whole input (headroom 0).
//
// The maximum a brick can receive is taken per frequency band: a rotation
// keeps |f|, so the pixels that reach a brick have radii |f| within the range
// the brick's coordinate box spans, and only the bands of that range count.
// When those bands' maxima agree within a factor kFlatRatio (a flat spectrum,
// or a brick far from DC), each contribution is rounded to a 32-bit quantum
// of 2^(e + headroom - 30), at most ~1e-6 of the smallest band maximum of the
// brick.  Otherwise (a brick spanning a steep part of the spectrum, such as
// the ones next to DC, whose voxels at shell 15 would be quantised against
// the DC value) the FP32 product w v is converted exactly to a 64-bit integer
// at scale 2^(62 - sumbits - headroom - e), sumbits bounding the bits of a
// voxel's sum, a quantum of ~1e-11 of the maximum.  Either way small
// high-frequency values keep their FP32 relative precision (per-shell
// relative error ~1e-7 for a spectrum falling 1e4 from DC to Nyquist, where
// one 32-bit scale per call gave 1e-3).
constexpr int kBands = 64;            // bands of |f| in cycles/pixel over [0, kBandTop)
constexpr float kFlatRatio = 4.0f;     // band maxima within this factor: 32-bit contributions
constexpr float kBandTop = 0.75f;     // above the largest slice radius sqrt(1/2)

__device__ __forceinline__ int band_of(float rho) {
  return min(kBands - 1, max(0, (int)(rho * ((float)kBands / kBandTop))));
}

struct ScaleArgs {
  const unsigned *maxbits;  // float bits of the exact max|v| of the input, kMaxSlots slots
  const unsigned *bandmax;  // float bits of the sampled max|v| per band (kBands words), or null
  unsigned *flag;           // set when a value exceeds the limit
  int headroom;             // bits of headroom above the sampled maxima
  int guard;                // 1: run only if *flag is set (the exact redo)
  // bits reserved for the sum of one voxel's 64-bit contributions: rotbits =
  // ceil(log2(rotations)) plus *slicebits, the largest per-slice bound of the
  // plane records (slice_count_bits)
  int rotbits;
  const unsigned *slicebits;
  // per brick, 1 if some slice plane can reach it, 0 if none does (then its
  // units return at once and its voxels are written by uncrossed_fill_kernel),
  // or null: every brick is processed
  const int *crossed;
};

// The block maxima are combined with atomicMax into kMaxSlots words kSlotStride
// words apart (separate cache lines, so the atomics do not queue on one
// address); the maximum is the largest slot.
constexpr int kMaxSlots = 32, kSlotStride = 32;
constexpr int kSetWords = kMaxSlots * kSlotStride;  // words of one set of slots
// scale words of a scatter call: sampled slots (unused), exact slots, the
// violation flag (on a line of its own), then the sampled band maxima
constexpr int kFlagWord = 2 * kSetWords, kBandWord = 2 * kSetWords + 32;
constexpr int kSliceWord = kBandWord + kBands;  // the largest per-slice count bits
constexpr int kScaleWords = kSliceWord + 1;

// Radius range (cycles/pixel) of the coordinate box [lo, hi] (array units):
// f_d = (c_d - b_d) / a_d with b = (bz, by, 0) and the array scales a.
__device__ __forceinline__ void box_radii(const SliceGeom &g, const float lo[3], const float hi[3], float &rmin,
                                          float &rmax) {
  const float b[3] = {(float)g.bz, (float)g.by, 0.0f};
  const float ia[3] = {(float)(1.0 / g.az), (float)(1.0 / g.ay), (float)(1.0 / g.ax)};
  float n2 = 0.0f, f2 = 0.0f;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float u = (lo[d] - b[d]) * ia[d], v = (hi[d] - b[d]) * ia[d];
    const float near = (u > 0.0f) ? u : ((v < 0.0f) ? -v : 0.0f);
    const float far = fmaxf(fabsf(u), fabsf(v));
    n2 += near * near;
    f2 += far * far;
  }
  rmin = sqrtf(n2);
  rmax = sqrtf(f2);
}

// The brick's scale, limit and contribution width (wide: 64-bit): from the
// sampled band maxima of the radii its cells span (one band of margin each
// side), or from the exact maximum of the input (always wide).
__device__ __forceinline__ void scale_of(const ScaleArgs &sa, const SliceGeom &g, const float lo[3],
                                         const float hi[3], float &scale, float &limit, bool &wide) {
  // one word per lane (kBands = 64 = the wave), one load in flight, then
  // wave reductions: the scale costs one memory latency per unit
  static_assert(kBands == 64 && kMaxSlots <= 64, "one band (or slot) per lane");
  const int lane = (int)(threadIdx.x & 63);
  unsigned mb = 0u;
  wide = true;
  if (sa.bandmax) {
    float rmin, rmax;
    box_radii(g, lo, hi, rmin, rmax);
    const int k0 = max(band_of(rmin * 0.9999f) - 1, 0), k1 = min(band_of(rmax * 1.0001f) + 1, kBands - 1);
    const unsigned b = sa.bandmax[lane];
    const bool in = lane >= k0 && lane <= k1;
    mb = in ? b : 0u;
    unsigned mn = (in && b) ? b : 0xffffffffu;
    for (int off = 32; off > 0; off >>= 1) {
      mb = max(mb, (unsigned)__shfl_xor((int)mb, off, 64));
      mn = min(mn, (unsigned)__shfl_xor((int)mn, off, 64));
    }
    wide = mb != 0u && __uint_as_float(mb) > kFlatRatio * __uint_as_float(mn);
  } else {
    mb = lane < kMaxSlots ? sa.maxbits[lane * kSlotStride] : 0u;
    for (int off = 32; off > 0; off >>= 1) mb = max(mb, (unsigned)__shfl_xor((int)mb, off, 64));
  }
  const float m = __uint_as_float(mb);
  int e = 0;
  if (m > 0.0f && m < 3.0e38f) frexpf(m, &e);
  e = max(e, -60);  // keeps the scales finite; values this small only lose relative precision
  const int sumbits = sa.rotbits + (int)*sa.slicebits;
  scale = ldexpf(1.0f, wide ? 62 - sumbits - sa.headroom - e : 30 - sa.headroom - e);
  limit = ldexpf(1.0f, e + sa.headroom);
}

constexpr float kNear = 2e-5f;            // FP32 guard band around integer coordinates

// Adds round(w * v) (half up) as a 64-bit fixed-point value (|w v| < 2^31).
__device__ __forceinline__ void lds_add_fixed(u64 *p, float w, float v) {
  atomicAdd(p, (u64)(long long)to_fixed32(w, v));
}

// Adds the FP32 product w v (v already multiplied by the power-of-two scale,
// which commutes with its rounding) as an exact 64-bit fixed-point value.
__device__ __forceinline__ void lds_add_wide(u64 *p, float w, float v) {
  atomicAdd(p, (u64)to_fixed64(__fmul_rn(w, v)));
}

// The exact canonical FP64 cell of pixel (i, j) under rotation r and Friedel
// half `neg`, for pixels whose FP32 tile coordinate lies within kNear of an
// integer: snap, faces and the Friedel half decide exactly as in every other
// kernel.  Returns false if the pixel belongs to the other half.
template <bool INSERT>
__device__ __forceinline__ bool exact_cell(const SliceGeom &g, const PlaneRec *__restrict__ pre, int r,
                                           bool neg, bool tie, int i, int j, int to0, int to1, int to2,
                                           int (&cell)[3], float (&frac)[3]) {
  const PlaneRec &p = pre[r];
  const double s = neg ? -1.0 : 1.0;
  const double fyd = canon_fy(g, i), jd = (double)j;
  const double cx1 = __fma_rn(jd, p.V[2], __dmul_rn(p.U[2], fyd));
  const bool ok = canon_conj<INSERT>(g, tie, cx1) == neg;
  double c[3];
  c[0] = snap(__fma_rn(jd, s * p.V[0], __fma_rn(s * p.U[0], fyd, g.bz)));
  c[1] = snap(__fma_rn(jd, s * p.V[1], __fma_rn(s * p.U[1], fyd, g.by)));
  c[2] = snap(s * cx1);
  const int top[3] = {g.vD - 1, g.vH - 1, g.vW - 1};
  const int to[3] = {to0, to1, to2};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double fl = floor(c[d]);
    int k = (int)fl;
    double t = c[d] - fl;
    if (k == top[d] && t == 0.0) {  // c = N-1 exactly: cell N-2 with fraction 1
      k = top[d] - 1;
      t = 1.0;
    }
    cell[d] = k - to[d];
    frac[d] = (float)t;
  }
  return ok;
}

// Decoded brick geometry shared by the scatter flush and combine kernels.
struct BrickBox {
  int lo[3], ext[3];  // first cell and local index of the last cell per axis
  bool last[3];       // the brick holds cell N-2: its upper voxel layer is the volume face
};

template <int BZ, int BY, int BX>
__device__ __forceinline__ BrickBox brick_box(const SliceGeom &g, const BrickGrid &bg, int brick) {
  BrickBox b;
  const int bz = brick / (bg.nby * bg.nbx);
  const int brem = brick - bz * bg.nby * bg.nbx;
  const int by = brem / bg.nbx, bx = brem - by * bg.nbx;
  const int B[3] = {BZ, BY, BX}, N[3] = {g.vD, g.vH, g.vW}, bi[3] = {bz, by, bx};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = bi[d] * B[d];
    const int ch = min(b.lo[d] + B[d] - 1, N[d] - 2);
    b.ext[d] = ch - b.lo[d];
    b.last[d] = ch == N[d] - 2;
  }
  return b;
}

// Writes a scatter brick's tile (NACC fixed-point planes of NT words):
// voxels it owns (local l in [0, ext], plus ext + 1 on axes where the brick
// is last) are converted and stored (the insertion's x = 0 plane into
// plane/wplane for the Friedel mirror pass); the upper shell (l = B on an
// axis where the brick is not last) belongs to the next brick and goes to
// the halo buffer as FP32 partial sums, converted with this brick's scale
// (unless write_shell is false: a brick no slice touched, whose shell the
// halo pass then skips).
template <int BZ, int BY, int BX, int MODE>
__device__ __forceinline__ void flush_voxel(const u64 *__restrict__ a, int astride, int v,
                                            const BrickBox &bb, const SliceGeom &g, int brick,
                                            float inv_v, int accumulate, f2v *__restrict__ out,
                                            float *__restrict__ wout, f2v *__restrict__ plane,
                                            float *__restrict__ wplane, float *__restrict__ halo,
                                            bool write_shell = true) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int SX = C::SX, SY = C::SY, NA = C::NACC, NSH = C::NSH;
  const int lz = v / (SY * SX), ly = (v / SX) % SY, lx = v % SX;
  if (lz > bb.ext[0] + 1 || ly > bb.ext[1] + 1 || lx > bb.ext[2] + 1) return;
  const bool shell = (lz == bb.ext[0] + 1 && !bb.last[0]) || (ly == bb.ext[1] + 1 && !bb.last[1]) ||
                     (lx == bb.ext[2] + 1 && !bb.last[2]);
  if (shell) {
    if (!write_shell) return;
    float *hb = halo + (int64_t)brick * NA * NSH;
    const int si = shell_index<BZ, BY, BX>(lz, ly, lx);
#pragma unroll
    for (int k = 0; k < NA; ++k) hb[k * NSH + si] = (float)(long long)a[k * astride] * (k < 2 ? inv_v : 1.0f / kWScale);
    return;
  }
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  const int gz = bb.lo[0] + lz, gy = bb.lo[1] + ly, gx = bb.lo[2] + lx;
  f2v val = {(float)(long long)a[0] * inv_v, (float)(long long)a[astride] * inv_v};
  if (MODE == kInsert) {
    const float w = (float)(long long)a[2 * astride] * (1.0f / kWScale);
    if (gx == 0) {  // the x = 0 plane goes through the Friedel mirror pass
      plane[(int64_t)gz * g.vH + gy] = val;
      wplane[(int64_t)gz * g.vH + gy] = w;
      return;
    }
    const int64_t e = gz * sz + gy * sy + gx;
    if (wout) wout[e] = accumulate ? wout[e] + w : w;  // no weights: a channel after the first
    out[e] = accumulate ? out[e] + val : val;
  } else {
    const int64_t e = gz * sz + gy * sy + gx;
    out[e] = accumulate ? out[e] + val : val;
  }
}

// acc: NACC planes of 64-bit words, or (split) the first two as planes of
// low and high words.
template <int BZ, int BY, int BX, int MODE>
__device__ void flush_tile(const u64 *__restrict__ acc, bool split, const SliceGeom &g, const BrickGrid &bg,
                           int brick, float vscale, int accumulate, f2v *__restrict__ out,
                           float *__restrict__ wout, f2v *__restrict__ plane,
                           float *__restrict__ wplane, float *__restrict__ halo, int tid, int nthreads,
                           bool write_shell) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int NT = C::NT, NA = C::NACC;
  const BrickBox bb = brick_box<BZ, BY, BX>(g, bg, brick);
  const float inv_v = 1.0f / vscale;
  const unsigned *a32 = reinterpret_cast<const unsigned *>(acc);
  for (int v = tid; v < NT; v += nthreads) {
    if (split) {
      u64 a[NA > 0 ? NA : 1];
#pragma unroll
      for (int k = 0; k < NA; ++k)
        a[k] = k < 2 ? split_value(a32[2 * k * NT + v], a32[(2 * k + 1) * NT + v]) : acc[k * NT + v];
      flush_voxel<BZ, BY, BX, MODE>(a, 1, v, bb, g, brick, inv_v, accumulate, out, wout, plane, wplane, halo,
                                    write_shell);
    } else {
      flush_voxel<BZ, BY, BX, MODE>(acc + v, NT, v, bb, g, brick, inv_v, accumulate, out, wout, plane,
                                    wplane, halo, write_shell);
    }
  }
}

// ---------------------------------------------------------------------------
// Gather zero pass: writes 0 to every pixel that the brick gather does not
// write, i.e. whose canonical coordinate (Friedel half included) lies
// outside the volume or, with fftfreq_max, outside the disc.  One lane per
// slice row computes, for each Friedel half, the interval of columns whose
// canonical coordinates are inside the volume by a margin of 1e-7 voxel,
// whose half decision is certain and (with fftfreq_max) whose disc test is
// certain, and the wider interval of columns that are possibly inside; the
// row's other columns (at most three runs) get a zero, except that a column
// in a possible interval is first tested with the canonical coordinate, as
// the pixel gather evaluates it.  The pass therefore writes exactly the
// pixels no brick writes, and may run before, after or between them.
// ---------------------------------------------------------------------------
struct ZeroRow {
  long long base;
  int a1, n1, a2, n2, a3, n3;  // three runs of columns [a, a + n)
  int r, i;                    // rotation and slice row
  int p[4];                    // columns possibly inside, per Friedel half: [p0, p1], [p2, p3]
};

__device__ __forceinline__ void lin_interval(double A, double B, double iB, double lo, double hi,
                                             double &jlo, double &jhi) {
  // lo <= A + j B <= hi, iB = 1 / B (the bounds carry a 1e-7 margin, far
  // above the rounding of the reciprocal)
  if (fabs(B) > 1e-30) {
    const double t1 = (lo - A) * iB, t2 = (hi - A) * iB;
    jlo = fmax(jlo, fmin(t1, t2));
    jhi = fmin(jhi, fmax(t1, t2));
  } else if (A < lo || A > hi) {
    jhi = -1e300;
  }
}

// Rows [block * NWZ * 64, (block + 1) * NWZ * 64) of the stack, one per lane
// of the block's NWZ waves, with per-wave LDS tables s_zr (row records) and
// s_mark (marks).
template <int NWZ>
__device__ __forceinline__ void zero_rows(const SliceGeom &g, const float *__restrict__ rot, int64_t n_rows,
                                          f2v *__restrict__ out, int64_t block, int n_chan,
                                          ZeroRow (*s_zr)[64], unsigned (*s_mark)[64]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t t = (block * NWZ + wave) * 64 + lane;
  int total = 0;
  if (t < n_rows) {
    const int64_t r = t / g.H;
    const int i = (int)(t - r * g.H);
    const Rot rr = load_rot(rot + r * 9, g);
    const double fy = canon_fy(g, i);
    const double m = 1e-7;
    double iV[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) iV[d] = fabs(rr.V[d]) > 1e-30 ? 1.0 / rr.V[d] : 0.0;
    const double top[3] = {(double)(g.vD - 1), (double)(g.vH - 1), (double)(g.vW - 1)};
    int a[2], b[2], pa[2], pb[2];
#pragma unroll
    for (int sh = 0; sh < 2; ++sh) {
      const double s = sh ? -1.0 : 1.0;
      const double ux = __dmul_rn(rr.U[2], fy);
      const double C0 = __fma_rn(s * rr.U[0], fy, g.bz), C1 = __fma_rn(s * rr.U[1], fy, g.by);
      // e = +1: certainly inside (margins inward); e = -1: possibly inside
#pragma unroll
      for (int ce = 0; ce < 2; ++ce) {
        const double e = ce ? -1.0 : 1.0;
        double jlo = 0.0, jhi = (double)(g.Wr - 1);
        // the Friedel half: the extraction decision flips within tolx of x' = 0
        lin_interval(s * ux, s * rr.V[2], s * iV[2], e * (g.tolx + m), 1e300, jlo, jhi);
        lin_interval(C0, s * rr.V[0], s * iV[0], e * m, top[0] - e * m, jlo, jhi);
        lin_interval(C1, s * rr.V[1], s * iV[1], e * m, top[1] - e * m, jlo, jhi);
        lin_interval(s * ux, s * rr.V[2], s * iV[2], e * m, top[2] - e * m, jlo, jhi);
        if (g.use_mask) {
          // in_disc (float32) holds wherever the exact radius is 1e-6 below
          // fmax, and fails wherever it is 1e-6 above
          const double fm = (double)g.fmax * (1.0 - e * 1e-6);
          const double fyv = (double)(i - g.half_H) * (double)g.inv_H32;
          const double rem = fm * fm - fyv * fyv;
          if (rem < 0.0 || fm < 0.0) jhi = -1e300;
          else jhi = fmin(jhi, sqrt(rem) / (double)g.inv_W32 - e * 1e-6);
        }
        int lo = 0, hi = -1;
        if (jlo <= jhi) {
          lo = max((int)ceil(jlo + e * 1e-7), 0);
          hi = min((int)floor(jhi - e * 1e-7), g.Wr - 1);
        }
        if (ce == 0) {
          a[sh] = lo;
          b[sh] = hi;
        } else {
          pa[sh] = lo;
          pb[sh] = hi;
        }
      }
    }
    // sort the two certain intervals, drop empty ones, merge overlaps
    int a1 = a[0], b1 = b[0], a2 = a[1], b2 = b[1];
    if (a1 > b1) { a1 = a2; b1 = b2; a2 = 0; b2 = -1; }
    if (a2 <= b2 && a2 < a1) { int ta = a1, tb = b1; a1 = a2; b1 = b2; a2 = ta; b2 = tb; }
    if (a2 <= b2 && a2 <= b1 + 1) { b1 = max(b1, b2); a2 = 0; b2 = -1; }
    ZeroRow z;
    z.base = (long long)r * g.ostride + (long long)i * g.Wr;
    z.r = (int)r;
    z.i = i;
    z.p[0] = pa[0];
    z.p[1] = pb[0];
    z.p[2] = pa[1];
    z.p[3] = pb[1];
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
  // the rows' zero pixels 64 at a time, consecutive lanes on consecutive
  // pixels of a row (coalesced stores); a lane finds its row from marks that
  // the rows write at their first pixel of the step, carried to the pixels
  // that follow by a running maximum over the wave
  const int incl = wave_incl_scan(total);
  const int excl = incl - total;
  const int total_px = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(incl, 63));
  unsigned *marks = s_mark[wave];
  for (int pb = 0; pb < total_px; pb += 64) {
    const unsigned long long hb = __ballot(excl <= pb && pb < incl);
    const int k0 = (int)__builtin_ctzll(hb);
    __builtin_amdgcn_wave_barrier();  // the previous step's marks have been read
    marks[lane] = lane == 0 ? (unsigned)(k0 + 1) : 0u;
    if (total > 0 && excl > pb && excl < pb + 64) marks[excl - pb] = (unsigned)(lane + 1);
    __builtin_amdgcn_wave_barrier();
    const int k = (int)wave_incl_max(marks[lane]) - 1;
    const int ek = __shfl(excl, k, 64);  // (every lane active)
    const int p = pb + lane;
    if (p < total_px) {
      const ZeroRow z = s_zr[wave][k];
      const int qq = p - ek;
      int j;
      if (qq < z.n1) j = z.a1 + qq;
      else if (qq < z.n1 + z.n2) j = z.a2 + (qq - z.n1);
      else j = z.a3 + (qq - z.n1 - z.n2);
      bool zero = true;
      if ((j >= z.p[0] && j <= z.p[1]) || (j >= z.p[2] && j <= z.p[3])) {
        // possibly inside: the canonical test of the pixel gather (a few
        // columns at the ends of the certain intervals)
        const Rot rj = load_rot(rot + (int64_t)z.r * 9, g);
        zero = !(in_disc(g, z.i, j) && inside(g, slice_coord<false>(g, rj, z.i, j)));
      }
      if (zero)
        for (int ch = 0; ch < n_chan; ++ch) out[z.base + (int64_t)ch * g.P + j] = f2v{0.0f, 0.0f};
    }
  }
}

__device__ __forceinline__ void gather_zero_body(const SliceGeom &g, const float *__restrict__ rot,
                                                 int64_t n_rows, f2v *__restrict__ out, int64_t block,
                                                 int n_chan = 1) {
  __shared__ ZeroRow s_zr[4][64];
  __shared__ unsigned s_mark[4][64];
  zero_rows<4>(g, rot, n_rows, out, block, n_chan, s_zr, s_mark);
}

// A row block of the zero pass run by a block of the fused gather launch
// (rot == nullptr: the block runs its brick unit instead).
struct ZeroArgs {
  const float *rot;
  int64_t n_rows;  // slice rows of the stack
  int64_t block;   // row block: rows [block * 64 * waves, (block + 1) * 64 * waves)
};

// ---------------------------------------------------------------------------
// The brick kernel (gather, adjoint or insertion)
// ---------------------------------------------------------------------------
// One unit (a brick and a range of rotations) of the gather, the adjoint or
// the insertion: shared by the one-call kernel and the batched persistent one.
// A gather block may run a row block of the zero pass instead (za.rot set).
template <int BZ, int BY, int BX, int MODE, bool MASK, int NCH = 1>
__device__ __forceinline__ void brick_body(
    const f2v *__restrict__ src, const SliceGeom &g, const BrickGrid &bg, const PlaneRec *__restrict__ pre,
    const PlaneTest *__restrict__ ptest, const Unit &un, int n_rot, f2v *__restrict__ out,
    float *__restrict__ wout, f2v *__restrict__ plane, float *__restrict__ wplane,
    u64 *__restrict__ segbuf, float *__restrict__ halo, const ScaleArgs &sa, int accumulate,
    int *__restrict__ touched, const ZeroArgs &za = ZeroArgs{nullptr, 0, 0}) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int NW = C::NW, CH = C::CH, SX = C::SX, SY = C::SY, NT = C::NT;
  constexpr bool SCATTER = MODE != kGather;
  // 32-bit contributions of the value (re, im) go to split accumulators;
  // the insertion's weights stay 64-bit words
  constexpr bool SPLIT = SCATTER;
  static_assert(CH >= 1 && CH <= 4, "chunks of 1 to 4 pixels");
  static_assert(NCH == 1 || (NCH == 2 && MODE == kGather), "two channels: gather only");
  // rotations per item pass (two halves each): 32 where the LDS of two
  // blocks per CU has room for 64 item records per wave, else 16
  constexpr int kTake = (MODE == kGather && NCH == 1) ? 32 : 16;
  // gather: NCH x NT complex voxels; scatter: NACC fixed-point planes of NT words
  constexpr int NWORDS = SCATTER ? C::NACC * NT : NCH * NT;
  // one LDS block, the tile first: the trilinear corner offsets then fit the
  // offset fields of the LDS instructions
  struct Smem {
    u64 tile[NWORDS];
    ItemRec item[NW][2 * kTake];
    // (rotation, Friedel half) entries whose plane meets the brick, pending:
    // rotation | 2^31 for the half s = -1
    int cross[NW][2 * kTake + 128];
    RowRec rec[NW][64];  // the rows of the current row pass
    // chunk step marks: row lane << 26 | first chunk of that row, at the
    // step slot where the row's chunks start
    unsigned mark[NW][64];
  };
  __shared__ Smem sm;
  u64 *s_tile = sm.tile;

  if (MODE == kGather && za.rot) {
    // the zero rows' tables alias the start of the block's LDS
    constexpr size_t kZeroTab = (size_t)NW * 64 * sizeof(ZeroRow);
    static_assert(sizeof(Smem) >= kZeroTab + (size_t)NW * 64 * sizeof(unsigned), "zero tables fit the LDS");
    char *base = reinterpret_cast<char *>(&sm);
    zero_rows<NW>(g, za.rot, za.n_rows, out, za.block, NCH, reinterpret_cast<ZeroRow(*)[64]>(base),
                  reinterpret_cast<unsigned(*)[64]>(base + kZeroTab));
    return;
  }
  if (SCATTER && sa.guard && *sa.flag == 0u) return;  // exact redo not needed
  const int brick = un.brick;
  if (SCATTER && sa.crossed && !sa.crossed[brick]) return;  // no plane reaches it
  const int seg = un.seg >> 16, nseg_raw = un.seg & 0xffff;
  const int nseg = max(nseg_raw, 1);  // an empty brick (no slice reaches it) only writes zeros
  const int bz = brick / (bg.nby * bg.nbx);
  const int brem = brick - bz * bg.nby * bg.nbx;
  const int by = brem / bg.nbx, bx = brem - by * bg.nbx;
  const int lo0 = bz * BZ, lo1 = by * BY, lo2 = bx * BX;
  // the brick's cells lo..ch (cells run 0..N-2 per axis)
  const int ch0 = min(lo0 + BZ - 1, g.vD - 2), ch1 = min(lo1 + BY - 1, g.vH - 2),
            ch2 = min(lo2 + BX - 1, g.vW - 2);
  const unsigned ext0 = (unsigned)(ch0 - lo0), ext1 = (unsigned)(ch1 - lo1), ext2 = (unsigned)(ch2 - lo2);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  float vscale = 1.0f, vlimit = 3.0e38f;
  bool wide = false;  // 64-bit contributions (uniform over the block)
  if (SCATTER) {
    const float blo[3] = {(float)lo0, (float)lo1, (float)lo2};
    const float bhi[3] = {(float)(ch0 + 1), (float)(ch1 + 1), (float)(ch2 + 1)};
    scale_of(sa, g, blo, bhi, vscale, vlimit, wide);
  }
  bool bad = false;  // a value beyond the fixed-point limit was seen

  if (!SCATTER) {
    f2v *tv = reinterpret_cast<f2v *>(s_tile);
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch)
      for (int v = tid; v < NT; v += C::T) {
        const int z = v / (SY * SX), y = (v / SX) % SY, x = v % SX;
        const int gz = lo0 + z, gy = lo1 + y, gx = lo2 + x;
        f2v val = {0.0f, 0.0f};
        if (gz < g.vD && gy < g.vH && gx < g.vW) val = src[ch * g.vstride + ((int64_t)gz * g.vH + gy) * g.vW + gx];
        tv[ch * NT + v] = val;
      }
  } else if (wide || !SPLIT) {
    for (int v = tid; v < NWORDS; v += C::T) s_tile[v] = 0ull;
  } else {  // split accumulators (re, im): low words kLoBias, high words 0; weights 0
    unsigned *t32 = reinterpret_cast<unsigned *>(s_tile);
    for (int v = tid; v < 2 * NWORDS; v += C::T) t32[v] = (v < 4 * NT && !((v / NT) & 1)) ? kLoBias : 0u;
  }
  __syncthreads();

  // brick constants: uniform over the block, held in scalar registers
  const double Lb[3] = {uniform_d((double)lo0 - kBoxMargin), uniform_d((double)lo1 - kBoxMargin),
                        uniform_d((double)lo2 - kBoxMargin)};
  const double Ub[3] = {uniform_d((double)(ch0 + 1) + kBoxMargin), uniform_d((double)(ch1 + 1) + kBoxMargin),
                        uniform_d((double)(ch2 + 1) + kBoxMargin)};
  // origin of the FP32 item frames: the tile centre, which halves the FP32
  // rounding of the coordinates against the tile corner
  constexpr int HZ = BZ / 2, HY = BY / 2, HX = BX / 2;
  const double tor[3] = {uniform_d((double)(lo0 + HZ)), uniform_d((double)(lo1 + HY)),
                         uniform_d((double)(lo2 + HX))};
  const double beta[3] = {g.bz, g.by, 0.0};
  double ctr[3], half[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ctr[d] = uniform_d(0.5 * (Lb[d] + Ub[d]));
    half[d] = uniform_d(0.5 * (Ub[d] - Lb[d]));
  }
  const double mmin = uniform_d((double)(-g.half_H)), mmax = uniform_d((double)(g.H - 1 - g.half_H));

  // this wave's share of the unit's rotations
  const int64_t ub = (int64_t)n_rot * seg / nseg, ue = (int64_t)n_rot * (seg + 1) / nseg;
  const int r_beg = (int)(ub + (ue - ub) * wave / NW);
  const int r_end = (int)(ub + (ue - ub) * (wave + 1) / NW);

  ItemRec *items = sm.item[wave];
  int *cross = sm.cross[wave];
  RowRec *recs = sm.rec[wave];
  unsigned *marks = sm.mark[wave];
  const bool have_cells = nseg_raw > 0 && ch0 >= lo0 && ch1 >= lo1 && ch2 >= lo2;
  const double cb0 = uniform_d(ctr[0] - beta[0]), cb1 = uniform_d(ctr[1] - beta[1]),
               cb2 = uniform_d(ctr[2] - beta[2]);
  // candidate box of the row phase in frame coordinates, widened well beyond
  // the FP32 error
  const float Lt[3] = {(float)(-HZ) - kRowMargin, (float)(-HY) - kRowMargin, (float)(-HX) - kRowMargin};
  // largest cell corner of the brick in frame coordinates
  const float fext0 = (float)((int)ext0 - HZ), fext1 = (float)((int)ext1 - HY), fext2 = (float)((int)ext2 - HX);
  const float Ut[3] = {(float)((int)ext0 + 1 - HZ) + kRowMargin, (float)((int)ext1 + 1 - HY) + kRowMargin,
                       (float)((int)ext2 + 1 - HX) + kRowMargin};

  // FP32 plane test constants (scalar registers); the margin covers the FP32
  // rounding of the test with a wide safety factor (extra candidates only
  // cost an empty item)
  const float cf0 = uniform_f((float)cb0), cf1 = uniform_f((float)cb1), cf2 = uniform_f((float)cb2);
  const float hf0 = uniform_f((float)half[0]), hf1 = uniform_f((float)half[1]), hf2 = uniform_f((float)half[2]);
  const float tmarg = uniform_f(1e-3f + 4e-6f * (fabsf(cf0) + fabsf(cf1) + fabsf(cf2) + hf0 + hf1 + hf2));
  // the box relative to beta (FP32), for the ranges of the box's section
  const float flo[3] = {uniform_f((float)(Lb[0] - beta[0])), uniform_f((float)(Lb[1] - beta[1])),
                        uniform_f((float)(Lb[2] - beta[2]))};
  const float fhi[3] = {uniform_f((float)(Ub[0] - beta[0])), uniform_f((float)(Ub[1] - beta[1])),
                        uniform_f((float)(Ub[2] - beta[2]))};
  const float fbig = uniform_f(fmaxf(fabsf(flo[0]), fabsf(fhi[0])) + fmaxf(fabsf(flo[1]), fabsf(fhi[1])) +
                               fmaxf(fabsf(flo[2]), fabsf(fhi[2])) + 1.0f);

  int rd = r_beg, ncross = 0;
  bool had_items = false;  // this wave listed a slice item (the tile may have been written)
  while (have_cells) {
    // (1a) plane test, 64 rotations per pass, up to 4 passes whose data are
    // loaded at once: a rotation whose plane meets the brick's coordinate box
    // is listed once for each Friedel half s whose columns can reach the box,
    // j = (x - beta) . s B* >= 0 somewhere on it (at most 2 kTake + 127
    // entries pending).  A brick off the line j = 0 of a plane receives one
    // half only, so the item lanes below are spent on items that exist.
    if (ncross < 2 * kTake && rd < r_end) {
      // all 4 passes are tested unconditionally (so that their loads are in
      // flight together); passes are then listed while the list has room and
      // the rest are tested again in the next round
      constexpr int kPass = 4;
      unsigned long long hp[kPass], hn[kPass];
      PlaneTest pt[kPass];
#pragma unroll
      for (int u = 0; u < kPass; ++u) {
        const int r = rd + u * 64 + lane;
        pt[u] = ptest[min(r, r_end - 1)];
      }
      __asm__ volatile("" ::: "memory");  // keep the loads ahead of their use
#pragma unroll
      for (int u = 0; u < kPass; ++u) {
        const int r = rd + u * 64 + lane;
        const float4 q = pt[u].n, b = pt[u].b;
        const float dist = q.x * cf0 + q.y * cf1 + q.z * cf2;
        const float rad = fabsf(q.x) * hf0 + fabsf(q.y) * hf1 + fabsf(q.z) * hf2;
        const bool full = q.w < 0.0f;
        const bool hit = r < r_end && (full | (fabsf(dist) <= rad + tmarg));
        // j over the box for s = +1 spans cj -+ hj; half s reaches j >= 0
        // when s cj + hj >= 0 (FP32, widened far beyond its rounding; an
        // extra entry only costs an empty item)
        const float cj = b.x * cf0 + b.y * cf1 + b.z * cf2;
        const float hj = fabsf(b.x) * hf0 + fabsf(b.y) * hf1 + fabsf(b.z) * hf2;
        const float jm = 1e-2f + 1e-5f * (fabsf(b.x * cf0) + fabsf(b.y * cf1) + fabsf(b.z * cf2) + hj);
        hp[u] = __ballot(hit && (full || cj >= -hj - jm));
        hn[u] = __ballot(hit && (full || cj <= hj + jm));
      }
#pragma unroll
      for (int u = 0; u < kPass; ++u) {
        if (ncross < 2 * kTake && rd < r_end) {
          const bool lp = (hp[u] >> lane) & 1ull, ln = (hn[u] >> lane) & 1ull;
          const int at = ncross + rank_below(hp[u]) + rank_below(hn[u]);
          if (lp) cross[at] = rd + lane;
          if (ln) cross[at + (lp ? 1 : 0)] = (int)((unsigned)(rd + lane) | 0x80000000u);
          ncross += __popcll(hp[u]) + __popcll(hn[u]);
          rd += 64;
        }
      }
    }
    if (ncross == 0) break;
    __builtin_amdgcn_wave_barrier();
    // (1b) items: lane c handles listed entry c (a rotation and a Friedel
    // half).  Its candidate rows and columns are those of the box: the ranges
    // of m = (x - beta) . s As and j = (x - beta) . s Bs over the box (a
    // superset of the box's section by the plane; rows without a pixel in
    // the box come out empty).  The item's FP32 frame is referred to the pixel
    // nearest the projection of the tile centre.
    const int take = min(ncross, 2 * kTake);
    bool h = false;
    ItemRec it;
    if (lane < take) {
      const int ent = cross[lane];
      const int r = ent & 0x7fffffff;
      const bool ineg = ent < 0;
      const PlaneRec &p = pre[r];
      const double s = ineg ? -1.0 : 1.0;
      // every field is used before the branch below, so the loads are issued together
      double A[3], B[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        A[d] = s * p.U[d] * g.inv_H;
        B[d] = s * p.V[d];
      }
      const double sU0 = s * p.U[0], sU1 = s * p.U[1], U2 = p.U[2];
      const unsigned rtie = (unsigned)p.tie;
      double cm = 0.0, hm = 0.0, cj = 0.0, hj = 0.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double cbd = d == 0 ? cb0 : (d == 1 ? cb1 : cb2);
        cm += cbd * p.As[d];
        hm += half[d] * fabs(p.As[d]);
        cj += cbd * p.Bs[d];
        hj += half[d] * fabs(p.Bs[d]);
      }
      cm *= s;
      cj *= s;
      double mlo = cm - hm - 1e-6, mhi = cm + hm + 1e-6;
      double jlo = cj - hj - 1e-6, jhi = cj + hj + 1e-6;
      if (p.full) {
        mlo = mmin;
        mhi = mmax;
        jlo = 0.0;
        jhi = (double)(g.Wr - 1);
      }
      int m0 = (int)ceil(fmax(mlo, mmin)), m1 = (int)floor(fmin(mhi, mmax));
      const double jc0 = ceil(fmax(jlo, 0.0)), jc1 = floor(fmin(jhi, (double)(g.Wr - 1)));
      h = m0 <= m1 && jc0 <= jc1;
      if (h && !p.full) {
        // tighter: the ranges of m and j over the box's section by the plane,
        // min a.y subject to y in the box (relative to beta) and n.y = 0, by
        // its Lagrange dual max_lam sum_d min((a_d - lam n_d) l_d, (a_d - lam n_d) u_d),
        // whose maximum lies at a breakpoint lam = a_k / n_k (FP32, widened
        // well beyond its rounding)
        const float nz = (float)(A[1] * B[2] - A[2] * B[1]), ny = (float)(A[2] * B[0] - A[0] * B[2]),
                    nx = (float)(A[0] * B[1] - A[1] * B[0]);
        const float nmax = fmaxf(fmaxf(fabsf(nz), fabsf(ny)), fabsf(nx));
        const float fa[3] = {(float)(s * p.As[0]), (float)(s * p.As[1]), (float)(s * p.As[2])};
        const float fb[3] = {(float)(s * p.Bs[0]), (float)(s * p.Bs[1]), (float)(s * p.Bs[2])};
        const float nn[3] = {nz, ny, nx};
        float lo_m = -3.0e38f, lo_j = -3.0e38f, hi_m = -3.0e38f, hi_j = -3.0e38f;  // hi: of -a
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          if (!(fabsf(nn[k]) > 1e-6f * nmax)) continue;
          const float rk = 1.0f / nn[k];
          const float lm = fa[k] * rk, lj = fb[k] * rk;
          float gm = 0.0f, gj = 0.0f, gm2 = 0.0f, gj2 = 0.0f;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const float cm = fa[d] - lm * nn[d], cj = fb[d] - lj * nn[d];
            gm += fminf(cm * flo[d], cm * fhi[d]);
            gm2 += fminf(-cm * flo[d], -cm * fhi[d]);
            gj += fminf(cj * flo[d], cj * fhi[d]);
            gj2 += fminf(-cj * flo[d], -cj * fhi[d]);
          }
          lo_m = fmaxf(lo_m, gm);
          hi_m = fmaxf(hi_m, gm2);
          lo_j = fmaxf(lo_j, gj);
          hi_j = fmaxf(hi_j, gj2);
        }
        if (lo_m > -1.0e38f) {
          const float mg = 1e-2f + 1e-5f * fbig * (fabsf(fa[0]) + fabsf(fa[1]) + fabsf(fa[2]));
          const float jg = 1e-2f + 1e-5f * fbig * (fabsf(fb[0]) + fabsf(fb[1]) + fabsf(fb[2]));
          m0 = max(m0, (int)ceilf(fmaxf(lo_m - mg, -1.0e9f)));
          m1 = min(m1, (int)floorf(fminf(-hi_m + mg, 1.0e9f)));
          h = m0 <= m1 && ceilf(fmaxf(lo_j - jg, 0.0f)) <= floorf(fminf(-hi_j + jg, (float)(g.Wr - 1)));
        }
      }
      if (h) {
        const int mref = min(max((int)fmin(fmax(rint(cm), -1e9), 1e9), m0), m1);
        const int jref = (int)fmin(fmax(rint(cj), 0.0), (double)(g.Wr - 1));
        const int iref = mref + g.half_H;
        // the canonical coordinate of the reference pixel (as in every kernel)
        const double fy = canon_fy(g, iref), jd = (double)jref;
        const double c0 = __fma_rn(jd, B[0], __fma_rn(sU0, fy, g.bz));
        const double c1 = __fma_rn(jd, B[1], __fma_rn(sU1, fy, g.by));
        const double c2 = __fma_rn(jd, B[2], s * __dmul_rn(U2, fy));
        it.st[0] = (float)(c0 - tor[0]);
        it.st[1] = (float)(c1 - tor[1]);
        it.st[2] = (float)(c2 - tor[2]);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          it.sa[d] = (float)A[d];
          it.sv[d] = (float)B[d];
        }
        it.rows = (m0 + g.half_H) | ((m1 - m0 + 1) << 16);
        it.ref = ((unsigned)iref & 0xffffu) | ((unsigned)jref << 16);
        it.rnt = (unsigned)r | (rtie << 30) | ((unsigned)ineg << 31);
      }
    }
    {  // keep the rest of the list (up to 127 entries, 64 at a time)
      const int rest = ncross - take;
      for (int b0 = 0; b0 < rest; b0 += 64) {
        const int v = b0 + lane < rest ? cross[take + b0 + lane] : 0;
        __builtin_amdgcn_wave_barrier();
        if (b0 + lane < rest) cross[b0 + lane] = v;
        __builtin_amdgcn_wave_barrier();
      }
      ncross = rest;
    }
    const unsigned long long bh = __ballot(h);
    const int nitems = __popcll(bh);
    if (h) items[rank_below(bh)] = it;
    __builtin_amdgcn_wave_barrier();
    if (nitems == 0) continue;
    had_items = true;

    // (2) inclusive prefix of the items' row counts: item k's end in lane k
    const int ia = wave_incl_scan(lane < nitems ? (items[lane].rows >> 16) : 0);
    const int total_rows = __builtin_amdgcn_readlane(ia, 63);

    // Row passes of 64 rows from row rb, whose first cskip chunks an earlier
    // pass processed.  Each row's lane writes the row's record; the chunks of
    // CH pixels of all rows are then taken 64 at a time, consecutive lanes on
    // consecutive chunks.  A chunk finds its row from marks: the lane of each
    // row whose chunks start inside the step writes its index at that slot,
    // and a running maximum over the wave carries it to the following slots.
    // A pass takes whole 64-lane steps of chunks while rows remain after it
    // (scatters), and the next pass starts at the row of the first chunk
    // left, so that chunk steps run with every lane busy except at the end.
    int rb = 0, cskip = 0;
    while (rb < total_rows) {
      // item k of each lane's row t (ia[k-1] <= t < ia[k]): kf items end at
      // or before the pass's first row; each item ending inside the pass
      // marks the slot of its end (the first row of the next item), and a
      // running maximum over the wave carries the marks to the rows that
      // follow (item ends are strictly increasing: every item has a row).
      // The row's offset in its item comes from the previous item's end.
      const int t = rb + lane;
      const int kf = __popcll(__ballot(lane < nitems && ia <= rb));
      marks[lane] = 0u;
      __builtin_amdgcn_wave_barrier();
      if (lane < nitems && ia > rb && ia < rb + 64) marks[ia - rb] = (unsigned)lane + 1u;
      __builtin_amdgcn_wave_barrier();
      const int k = max(kf, (int)wave_incl_max(marks[lane]));
      const int kprev = __builtin_amdgcn_ds_bpermute(max(k - 1, 0) << 2, ia);
      const int excl = k > 0 ? kprev : 0;
      __builtin_amdgcn_wave_barrier();  // the chunk steps rewrite the marks
      // one lane per row: the candidate columns, from the item's FP32 frame
      int len = 0;
      if (t < total_rows) {
        const ItemRec &ir = items[k];
        const int i = (ir.rows & 0xffff) + (t - excl);
        const int iref = (int)(ir.ref & 0xffffu), jref = (int)(ir.ref >> 16);
        const float fdm = (float)(i - iref);
        float lo = (float)(-jref), hi = (float)(g.Wr - 1 - jref);
        float C[3], sv[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          C[d] = __fmaf_rn(fdm, ir.sa[d], ir.st[d]);  // at column jref
          sv[d] = ir.sv[d];
          if (fabsf(sv[d]) > 1e-20f) {
            const float iv = __builtin_amdgcn_rcpf(sv[d]);
            const float t1 = (Lt[d] - C[d]) * iv, t2 = (Ut[d] - C[d]) * iv;
            lo = fmaxf(lo, fminf(t1, t2));
            hi = fminf(hi, fmaxf(t1, t2));
          } else if (C[d] < Lt[d] || C[d] > Ut[d]) {
            hi = -1.0e9f;
          }
        }
        int jl = 0;
        if (lo <= hi) {
          jl = jref + (int)ceilf(lo - 1e-3f - 1e-6f * fabsf(lo));
          int jh = jref + (int)floorf(hi + 1e-3f + 1e-6f * fabsf(hi));
          jl = max(jl, 0);
          jh = min(jh, g.Wr - 1);
          if (MODE == kInsert && i < g.half_H) jl = max(jl, 1);  // x = 0, y < 0 is skipped
          len = max(0, jh - jl + 1);
        }
        if (lane == 0) {  // chunks of the first row processed by the previous pass
          jl += cskip * CH;
          len = max(0, len - cskip * CH);
        }
        if (len > 0) {
          RowRec rec;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            rec.rowc[d] = C[d];
            rec.sv[d] = sv[d];
          }
          rec.djl = (float)(jl - jref);
          rec.ijl = (unsigned)i | ((unsigned)jl << 16);
          rec.len = len;
          rec.rnt = ir.rnt;
          recs[lane] = rec;
        }
      }
      const int nch = (len + CH - 1) / CH;
      const int cincl = wave_incl_scan(nch);
      const int cexcl = cincl - nch;
      const int total_ch = __builtin_amdgcn_readfirstlane(__builtin_amdgcn_readlane(cincl, 63));
      // (the gather measured no faster with partial steps carried over)
      const int done = (!SCATTER || rb + 64 >= total_rows || total_ch < 64) ? total_ch : (total_ch & ~63);

      // (3) chunk steps
      for (int cb = 0; cb < done; cb += 64) {
        // the row holding chunk cb (every chunk index below total_ch has one)
        const unsigned long long hb = __ballot(cexcl <= cb && cb < cincl);
        const int k0 = (int)__builtin_ctzll(hb);
        const int cx0 = __builtin_amdgcn_readlane(cexcl, k0);
        __builtin_amdgcn_wave_barrier();  // the previous step's marks have been read
        marks[lane] = lane == 0 ? ((unsigned)k0 << 26) | (unsigned)cx0 : 0u;
        if (nch > 0 && cexcl > cb && cexcl < cb + 64) marks[cexcl - cb] = ((unsigned)lane << 26) | (unsigned)cexcl;
        __builtin_amdgcn_wave_barrier();
        const unsigned mk = wave_incl_max(marks[lane]);
        const int kr = (int)(mk >> 26);
        const int gc = cb + lane;
        const bool live = gc < done;
        const RowRec &row = recs[kr];
        // first pixel of the chunk in its row (lanes past the last chunk take
        // the row's first, so that every address below stays inside the row)
        const int q0 = live ? (gc - (int)(mk & 0x3ffffffu)) * CH : 0;
        const int cnt = live ? min(CH, row.len - q0) : 0;
        const int i = (int)(row.ijl & 0xffffu);
        const int j0 = (int)(row.ijl >> 16) + q0;
        const int rr = (int)(row.rnt & 0x3fffffffu);
        const bool neg = row.rnt >> 31;
        // the Friedel half's sign bit, for one XOR with the imaginary part
        const unsigned nbit = row.rnt & 0x80000000u;
        // flat index of the chunk's first pixel (i Wr + j0 < H Wr < 2^31);
        // scatter input rotation stride: g.ostride (P, or C P for one
        // channel of a multichannel stack)
        const int64_t base = (int64_t)((u64)(unsigned)rr * (u64)(SCATTER ? g.ostride : (int64_t)g.P) +
                                       (u64)(unsigned)(i * g.Wr + j0));
        // scatter: the chunk's values are loaded first, so that the load
        // overlaps the coordinate arithmetic (pixels q < cnt are in the row)
        f2v vv[CH];
        if (SCATTER) {
#pragma unroll
          for (int q = 0; q < CH; ++q) vv[q] = src[q < cnt ? base + q : base];
        }
        // FP32 frame coordinates of the chunk's pixels: cell corner fl (the
        // cell relative to the tile centre) and fraction t; the column
        // offset from jref is an integer, exact in FP32
        float rowc[3], sv3[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          rowc[d] = row.rowc[d];
          sv3[d] = row.sv[d];
        }
        const float djq = row.djl + (float)q0;
        float fl[CH][3], t3[CH][3];
        float tmin = 1.0f, tmax = 0.0f;
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          const float qf = djq + (float)q;
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const float l = __fmaf_rn(qf, sv3[d], rowc[d]);
            fl[q][d] = floorf(l);
            t3[q][d] = l - fl[q][d];
          }
          tmin = fminf(tmin, fminf(fminf(t3[q][0], t3[q][1]), t3[q][2]));
          tmax = fmaxf(tmax, fmaxf(fmaxf(t3[q][0], t3[q][1]), t3[q][2]));
        }
        bool own[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) own[q] = q < cnt;
        const bool anynear = cnt > 0 && (tmin < kNear || tmax > 1.0f - kNear);
        if (__ballot(anynear)) {  // rare, wave-uniform: exact FP64 cells near integers
          if (anynear) {
#pragma unroll
            for (int q = 0; q < CH; ++q) {
              if (!own[q]) continue;
              int cell[3];
              own[q] = exact_cell<MODE == kInsert>(g, pre, rr, neg, (row.rnt >> 30) & 1, i, j0 + q, lo0,
                                                   lo1, lo2, cell, t3[q]);
              fl[q][0] = (float)(cell[0] - HZ);
              fl[q][1] = (float)(cell[1] - HY);
              fl[q][2] = (float)(cell[2] - HX);
            }
          }
        }
        // owned: the cell lies in the brick; e = its tile index
        int e[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) {
          own[q] = own[q] & (__builtin_amdgcn_fmed3f(fl[q][0], (float)-HZ, fext0) == fl[q][0]) &
                   (__builtin_amdgcn_fmed3f(fl[q][1], (float)-HY, fext1) == fl[q][1]) &
                   (__builtin_amdgcn_fmed3f(fl[q][2], (float)-HX, fext2) == fl[q][2]);
          if (MASK) own[q] = own[q] && in_disc(g, i, j0 + q);
          const float ef = __fmaf_rn(fl[q][0], (float)(SY * SX), __fmaf_rn(fl[q][1], (float)SX, fl[q][2]));
          e[q] = own[q] ? (int)ef + (HZ * SY * SX + HY * SX + HX) : 0;
        }
        if (MODE == kGather) {
          // channel ch of the output: rotation stride g.ostride, channel stride g.P
          int64_t obase = base;
          if (NCH > 1)
            obase = (int64_t)((u64)(unsigned)rr * (u64)g.ostride + (u64)(unsigned)(i * g.Wr + j0));
#pragma unroll
          for (int ch = 0; ch < NCH; ++ch) {
            const f2v *tv = reinterpret_cast<const f2v *>(s_tile) + ch * NT;
            f2v cv[CH][8];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
              const f2v *b = tv + e[q];
              cv[q][0] = b[0];
              cv[q][1] = b[1];
              cv[q][2] = b[SX];
              cv[q][3] = b[SX + 1];
              cv[q][4] = b[SY * SX];
              cv[q][5] = b[SY * SX + 1];
              cv[q][6] = b[SY * SX + SX];
              cv[q][7] = b[SY * SX + SX + 1];
            }
#pragma unroll
            for (int q = 0; q < CH; ++q) {
              f2v val = trilerp(cv[q][0], cv[q][1], cv[q][2], cv[q][3], cv[q][4], cv[q][5], cv[q][6],
                                cv[q][7], t3[q][0], t3[q][1], t3[q][2]);
              val.y = __uint_as_float(__float_as_uint(val.y) ^ nbit);  // conj on the s = -1 half
              if (own[q]) out[obase + ch * g.P + q] = val;
            }
          }
        } else {
#pragma unroll
          for (int q = 0; q < CH; ++q) {
            if (!own[q]) continue;
            f2v v = vv[q];
            v.y = __uint_as_float(__float_as_uint(v.y) ^ nbit);
            bad = bad || !(fabsf(v.x) < vlimit && fabsf(v.y) < vlimit);
            v.x *= vscale;
            v.y *= vscale;
            const float tz = t3[q][0], ty = t3[q][1], tx = t3[q][2];
            const float wz[2] = {1.0f - tz, tz}, wy[2] = {1.0f - ty, ty}, wx[2] = {1.0f - tx, tx};
            if (wide) {
#pragma unroll
              for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy) {
                  const float wzy = wz[dz] * wy[dy];
#pragma unroll
                  for (int dx = 0; dx < 2; ++dx) {
                    const float w = wzy * wx[dx];
                    const int o = e[q] + dz * SY * SX + dy * SX + dx;
                    lds_add_wide(&s_tile[o], w, v.x);
                    lds_add_wide(&s_tile[NT + o], w, v.y);
                    if (MODE == kInsert && wout) lds_add_fixed(&s_tile[2 * NT + o], w, kWScale);
                  }
                }
            } else if (!SPLIT) {
#pragma unroll
              for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy) {
                  const float wzy = wz[dz] * wy[dy];
#pragma unroll
                  for (int dx = 0; dx < 2; ++dx) {
                    const float w = wzy * wx[dx];
                    const int o = e[q] + dz * SY * SX + dy * SX + dx;
                    lds_add_fixed(&s_tile[o], w, v.x);
                    lds_add_fixed(&s_tile[NT + o], w, v.y);
                    if (MODE == kInsert && wout) lds_add_fixed(&s_tile[2 * NT + o], w, kWScale);
                  }
                }
            } else {
              // split accumulators (re, im): low words first, then the rare carries
              unsigned *t32 = reinterpret_cast<unsigned *>(s_tile);
              unsigned old[8][2];
#pragma unroll
              for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                for (int dy = 0; dy < 2; ++dy) {
                  const float wzy = wz[dz] * wy[dy];
#pragma unroll
                  for (int dx = 0; dx < 2; ++dx) {
                    const int k = dz * 4 + dy * 2 + dx;
                    const float w = wzy * wx[dx];
                    const int o = e[q] + dz * SY * SX + dy * SX + dx;
                    old[k][0] = atomicAdd(&t32[o], (unsigned)to_fixed32(w, v.x));
                    old[k][1] = atomicAdd(&t32[2 * NT + o], (unsigned)to_fixed32(w, v.y));
                    if (MODE == kInsert && wout) lds_add_fixed(&s_tile[2 * NT + o], w, kWScale);
                  }
                }
              if (!lo_all_in_window<16>(&old[0][0])) {
                // the contributions again, from the same operands (the
                // empty asm keeps the compiler from holding the first
                // products in registers across the atomics)
                float rx = v.x, ry = v.y, rz = tz, rw = ty, rt = tx;
                asm volatile("" : "+v"(rx), "+v"(ry), "+v"(rz), "+v"(rw), "+v"(rt));
                const float uz[2] = {1.0f - rz, rz}, uy[2] = {1.0f - rw, rw}, ux[2] = {1.0f - rt, rt};
#pragma unroll
                for (int dz = 0; dz < 2; ++dz)
#pragma unroll
                  for (int dy = 0; dy < 2; ++dy) {
                    const float wzy = uz[dz] * uy[dy];
#pragma unroll
                    for (int dx = 0; dx < 2; ++dx) {
                      const int k = dz * 4 + dy * 2 + dx;
                      const float w = wzy * ux[dx];
                      const int o = e[q] + dz * SY * SX + dy * SX + dx;
                      const unsigned h0 = hi_step(old[k][0], to_fixed32(w, rx));
                      const unsigned h1 = hi_step(old[k][1], to_fixed32(w, ry));
                      if (h0) atomicAdd(&t32[NT + o], h0);
                      if (h1) atomicAdd(&t32[3 * NT + o], h1);
                    }
                  }
              }
            }
          }
        }
      }
      if (done < total_ch) {  // the next pass starts at the row holding chunk `done`
        const int kr = __popcll(__ballot(cexcl + nch <= done));
        cskip = __builtin_amdgcn_readfirstlane(done - __builtin_amdgcn_readlane(cexcl, kr) + (kr == 0 ? cskip : 0));
        rb = __builtin_amdgcn_readfirstlane(rb + kr);
      } else {
        rb += 64;
        cskip = 0;
      }
      __builtin_amdgcn_wave_barrier();  // the next row pass rewrites this wave's row table
    }
    __builtin_amdgcn_wave_barrier();  // the next discovery pass rewrites the item table
  }

  if (!SCATTER) return;
  if (__ballot(bad) && lane == 0) atomicOr(sa.flag, 1u);
  const bool any = __syncthreads_or((int)had_items) != 0;
  if (nseg > 1) {  // partial sums of a rotation segment: combined by split_combine_kernel
    u64 *dst = segbuf + ((int64_t)un.slot + seg) * (C::NACC * NT);
    const unsigned *t32 = reinterpret_cast<const unsigned *>(s_tile);
    for (int v = tid; v < C::NACC * NT; v += C::T) {
      const int a = v / NT, u = v - a * NT;
      dst[v] = (wide || !SPLIT || a >= 2) ? s_tile[v] : split_value(t32[2 * a * NT + u], t32[(2 * a + 1) * NT + u]);
    }
    return;
  }
  // touched[brick]: whether any slice item reached the tile.  An untouched
  // tile is zero: its shell is not written to the halo buffer (the halo pass
  // skips it) and, when accumulating, nothing at all is written.
  if (touched) {
    if (tid == 0) touched[brick] = any ? 1 : 0;
    if (!any && accumulate) return;
  }
  flush_tile<BZ, BY, BX, MODE>(s_tile, SPLIT && !wide, g, bg, brick, vscale, accumulate, out, wout, plane, wplane,
                               halo, tid, C::T, any || !touched);
}

template <int BZ, int BY, int BX, int MODE, bool MASK, int NCH = 1>
__global__ __launch_bounds__((Cfg<BZ, BY, BX, MODE>::T), 4) void brick_kernel(
    const f2v *__restrict__ src, SliceGeom g, BrickGrid bg, const PlaneRec *__restrict__ pre,
    const PlaneTest *__restrict__ ptest, const Unit *__restrict__ units, int n_rot,
    f2v *__restrict__ out, float *__restrict__ wout, f2v *__restrict__ plane,
    float *__restrict__ wplane, u64 *__restrict__ segbuf, float *__restrict__ halo, ScaleArgs sa,
    int accumulate, int *__restrict__ touched) {
  // one unit per block (a stride loop over units measured 4-25 % slower, and
  // triples the adjoint's register spills)
  const Unit un = units[blockIdx.x];
  brick_body<BZ, BY, BX, MODE, MASK, NCH>(src, g, bg, pre, ptest, un, n_rot, out, wout, plane, wplane, segbuf,
                                          halo, sa, accumulate, touched);
}

// The gather in one launch: the plan's U brick units and the zero pass's Z
// row blocks (64 rows per wave), the row blocks spread evenly over the grid
// so that their stores overlap the brick units' VALU-bound work.  Block b
// runs row block k(b) - 1 where k(b) = floor((b + 1) Z / (U + Z)) steps up
// at b, and otherwise unit b - k(b).  The zero pass writes exactly the pixels
// that no unit writes, so the order of the blocks does not matter.
template <int BZ, int BY, int BX, bool MASK, int NCH = 1>
__global__ __launch_bounds__((Cfg<BZ, BY, BX, kGather>::T), 4) void gather_kernel(
    const f2v *__restrict__ src, SliceGeom g, BrickGrid bg, const PlaneRec *__restrict__ pre,
    const PlaneTest *__restrict__ ptest, const Unit *__restrict__ units, int n_units, int n_rot,
    const float *__restrict__ rot, int n_zero, f2v *__restrict__ out) {
  const int64_t b = blockIdx.x, tot = (int64_t)n_units + n_zero;
  const int64_t k = (b + 1) * n_zero / tot;
  const bool zero = k > b * n_zero / tot;
  const ZeroArgs za = {zero ? rot : nullptr, (int64_t)n_rot * g.H, k - 1};
  const Unit un = zero ? Unit{0, 0, 0} : units[b - k];
  const ScaleArgs sa = {nullptr, nullptr, nullptr, 0, 0, 0, nullptr, nullptr};
  brick_body<BZ, BY, BX, kGather, MASK, NCH>(src, g, bg, pre, ptest, un, n_rot, out, nullptr, nullptr, nullptr,
                                             nullptr, nullptr, sa, 0, nullptr, za);
}

// Sums the rotation segments of a split brick (integer sums: the result does
// not depend on the order) and flushes the tile like an unsplit brick; one
// thread per tile voxel, blockIdx.x over the split bricks.
template <int BZ, int BY, int BX, int MODE>
__device__ __forceinline__ void split_body(
    const SliceGeom &g, const BrickGrid &bg, const SplitRec &sp, const u64 *__restrict__ segbuf,
    const ScaleArgs &sa, int accumulate, f2v *__restrict__ out, float *__restrict__ wout,
    f2v *__restrict__ plane, float *__restrict__ wplane, float *__restrict__ halo, int v,
    int *__restrict__ touched) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int NT = C::NT, NA = C::NACC;
  if (sa.guard && *sa.flag == 0u) return;
  if (sa.crossed && !sa.crossed[sp.brick]) return;
  // the scale first, with every lane of the wave (scale_of reduces over it)
  const BrickBox bb = brick_box<BZ, BY, BX>(g, bg, sp.brick);
  float vscale, vlimit;
  {
    const float blo[3] = {(float)bb.lo[0], (float)bb.lo[1], (float)bb.lo[2]};
    const float bhi[3] = {(float)(bb.lo[0] + bb.ext[0] + 1), (float)(bb.lo[1] + bb.ext[1] + 1),
                          (float)(bb.lo[2] + bb.ext[2] + 1)};
    bool wide;
    scale_of(sa, g, blo, bhi, vscale, vlimit, wide);
  }
  if (v >= NT) return;
  if (touched && v == 0) touched[sp.brick] = 1;  // split bricks are the busiest ones
  u64 acc[NA];
#pragma unroll
  for (int a = 0; a < NA; ++a) acc[a] = 0ull;
  for (int sgi = 0; sgi < sp.nseg; ++sgi) {
    const u64 *b = segbuf + ((int64_t)sp.slot + sgi) * (NA * NT);
#pragma unroll
    for (int a = 0; a < NA; ++a) acc[a] += b[a * NT + v];
  }
  flush_voxel<BZ, BY, BX, MODE>(acc, 1, v, bb, g, sp.brick, 1.0f / vscale, accumulate, out, wout, plane,
                                wplane, halo);
}

template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__(256) void split_combine_kernel(
    SliceGeom g, BrickGrid bg, const SplitRec *__restrict__ splits, const u64 *__restrict__ segbuf,
    ScaleArgs sa, int accumulate, f2v *__restrict__ out, float *__restrict__ wout,
    f2v *__restrict__ plane, float *__restrict__ wplane, float *__restrict__ halo, int *__restrict__ touched) {
  const SplitRec sp = splits[blockIdx.x];
  split_body<BZ, BY, BX, MODE>(g, bg, sp, segbuf, sa, accumulate, out, wout, plane, wplane, halo,
                               (int)(blockIdx.y * 256 + threadIdx.x), touched);
}

// Adds the halo partial sums of the lower neighbours to the voxels of a
// brick's lower shell (l = 0 on an axis where the brick has a neighbour
// below): a voxel on a face, edge or corner receives 1, 3 or 7 halos.
// With touched set, the halos of untouched neighbours (not written, zero) are
// skipped, and so is a voxel that receives none.
template <int BZ, int BY, int BX, int MODE>
__device__ __forceinline__ void halo_body(const SliceGeom &g, const BrickGrid &bg,
                                          const float *__restrict__ halo, const ScaleArgs &sa,
                                          f2v *__restrict__ out, float *__restrict__ wout,
                                          f2v *__restrict__ plane, float *__restrict__ wplane, int brick,
                                          int u, const int *__restrict__ touched) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  constexpr int SX = C::SX, SY = C::SY, NA = C::NACC, NSH = C::NSH;
  if (sa.guard && *sa.flag == 0u) return;
  if (u >= NSH) return;
  const BrickBox bb = brick_box<BZ, BY, BX>(g, bg, brick);
  const int B[3] = {BZ, BY, BX};
  const int bstride[3] = {bg.nby * bg.nbx, bg.nbx, 1};
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  {
    // lower shell in the same order as the upper one: Z face, Y face, X face
    int l[3];
    if (u < SY * SX) {
      l[0] = 0;
      l[1] = u / SX;
      l[2] = u % SX;
    } else if (u < SY * SX + BZ * SX) {
      const int w = u - SY * SX;
      l[0] = 1 + w / SX;
      l[1] = 0;
      l[2] = w % SX;
    } else {
      const int w = u - SY * SX - BZ * SX;
      l[0] = 1 + w / BY;
      l[1] = 1 + w % BY;
      l[2] = 0;
    }
    bool owned = true;
    int amask = 0;  // axes on which the voxel receives a lower neighbour's halo
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      owned = owned && l[d] <= bb.ext[d] + (bb.last[d] ? 1 : 0);
      if (l[d] == 0 && bb.lo[d] > 0) amask |= 1 << d;
    }
    if (!owned || amask == 0) return;
    // FP32 partial sums (each converted with its own brick's scale), added in
    // the fixed order of the loop below: the result does not depend on timing
    float acc[3] = {0.0f, 0.0f, 0.0f};
    int nread = 0;
    // the nonempty subsets of amask, unrolled so that every load is in flight at once
#pragma unroll
    for (int sbits = 1; sbits < 8; ++sbits) {
      if ((sbits & amask) != sbits) continue;
      int nb = brick, lp[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const bool in = (sbits >> d) & 1;
        nb -= in ? bstride[d] : 0;
        lp[d] = l[d] + (in ? B[d] : 0);
      }
      if (touched && !touched[nb]) continue;
      ++nread;
      const int si = shell_index<BZ, BY, BX>(lp[0], lp[1], lp[2]);
      const float *hb = halo + (int64_t)nb * NA * NSH;
#pragma unroll
      for (int a = 0; a < NA; ++a) acc[a] += hb[a * NSH + si];
    }
    if (touched && nread == 0) return;
    const int gz = bb.lo[0] + l[0], gy = bb.lo[1] + l[1], gx = bb.lo[2] + l[2];
    const f2v val = {acc[0], acc[1]};
    if (MODE == kInsert) {
      const float w = acc[2];
      if (gx == 0) {
        const int64_t pe = (int64_t)gz * g.vH + gy;
        plane[pe] += val;
        wplane[pe] += w;
        return;
      }
      const int64_t e = gz * sz + gy * sy + gx;
      if (wout) wout[e] += w;
      out[e] += val;
    } else {
      const int64_t e = gz * sz + gy * sy + gx;
      out[e] += val;
    }
  }
}

template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__(256) void halo_combine_kernel(SliceGeom g, BrickGrid bg,
                                                           const float *__restrict__ halo, ScaleArgs sa,
                                                           f2v *__restrict__ out,
                                                           float *__restrict__ wout,
                                                           f2v *__restrict__ plane,
                                                           float *__restrict__ wplane,
                                                           const int *__restrict__ touched) {
  const int n_bricks = bg.nbz * bg.nby * bg.nbx;
  for (int b = blockIdx.x; b < n_bricks; b += gridDim.x)
    halo_body<BZ, BY, BX, MODE>(g, bg, halo, sa, out, wout, plane, wplane, b,
                                (int)(blockIdx.y * 256 + threadIdx.x), touched);
}

// max(|re|, |im|) as float bits into *res, over chunks of 8192 consecutive
// values (64 KiB) taken every `stride_chunks` chunks: 1 reads everything, 64
// a 1/64 sample of contiguous blocks.  Four loads in flight per lane; one
// atomic per block, into slot blockIdx.x % kMaxSlots.
__device__ __forceinline__ void maxabs_body(const f2v *__restrict__ x, int64_t n, int64_t stride_chunks,
                                            const unsigned *guard, unsigned *__restrict__ res, int block,
                                            int nblocks) {
  __shared__ float s_m[4];
  if (guard && *guard == 0u) return;
  const int64_t lanes = (int64_t)nblocks * 256;
  float m = 0.0f;
  for (int64_t t0 = (int64_t)block * 256 + threadIdx.x;; t0 += 4 * lanes) {
    f2v v[4];
    bool any = false;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t t = t0 + u * lanes;
      const int64_t e = (t >> 13) * stride_chunks * 8192 + (t & 8191);
      v[u] = e < n ? x[e] : f2v{0.0f, 0.0f};
      any = any || e < n;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) m = fmaxf(m, fmaxf(fabsf(v[u].x), fabsf(v[u].y)));
    if (!any) break;
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
    atomicMax(res + (block % kMaxSlots) * kSlotStride,
              __float_as_uint(fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]))));
}

__global__ __launch_bounds__(256) void maxabs_chunks_kernel(const f2v *__restrict__ x, int64_t n,
                                                            int64_t stride_chunks,
                                                            const unsigned *guard,
                                                            unsigned *__restrict__ res) {
  maxabs_body(x, n, stride_chunks, guard, res, (int)blockIdx.x, (int)gridDim.x);
}

// Sampled per-band max(|re|, |im|) of a slice stack: every rot_stride-th
// slice, each pixel's value into the band of its radius |f|.  A thread takes
// pixels (consecutive threads, consecutive pixels) and walks the sampled
// slices for each, so a pixel's band is computed once; block maxima in LDS,
// then one atomic per band and block.
__device__ __forceinline__ void band_max_body(const f2v *__restrict__ x, const SliceGeom &g, int64_t n_rot,
                                              int64_t rot_stride, unsigned *__restrict__ bands, int block,
                                              int nblocks, int sgroup = 0, int ngroups = 1) {
  __shared__ unsigned s_b[kBands];
  for (int k = threadIdx.x; k < kBands; k += blockDim.x) s_b[k] = 0u;
  __syncthreads();
  const int64_t nsl = (n_rot + rot_stride - 1) / rot_stride;
  const int64_t step = rot_stride * g.ostride;
  const float ih = (float)g.inv_H, iw = (float)g.inv_W;
  for (int p = block * (int)blockDim.x + (int)threadIdx.x; p < g.P; p += nblocks * (int)blockDim.x) {
    int i, j;
    split_pixel(g, p, i, j);
    const float fy = (float)(i - g.half_H) * ih, fx = (float)j * iw;
    // this block's sampled slices: sgroup, sgroup + ngroups, ...
    float m = 0.0f;
    int64_t sl = sgroup;
    const int64_t gstep = (int64_t)ngroups * step;
    for (; sl + 7 * (int64_t)ngroups < nsl; sl += 8 * (int64_t)ngroups) {  // eight loads in flight
      f2v v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = x[sl * step + u * gstep + p];
#pragma unroll
      for (int u = 0; u < 8; ++u) m = fmaxf(m, fmaxf(fabsf(v[u].x), fabsf(v[u].y)));
    }
    for (; sl < nsl; sl += ngroups) {
      const f2v v = x[sl * step + p];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
    atomicMax(&s_b[band_of(sqrtf(fy * fy + fx * fx))], __float_as_uint(m));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kBands; k += blockDim.x)
    if (s_b[k]) atomicMax(bands + k, s_b[k]);
}

__global__ __launch_bounds__(256) void band_max_kernel(const f2v *__restrict__ x, SliceGeom g, int64_t n_rot,
                                                       int64_t rot_stride, unsigned *__restrict__ bands) {
  band_max_body(x, g, n_rot, rot_stride, bands, (int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y, (int)gridDim.y);
}

// Slices sampled for the band maxima: about 64 spread over the stack.
__host__ __device__ inline int64_t band_stride(int64_t n_rot) { return n_rot > 64 ? n_rot / 64 : 1; }

// Blocks of a band-maximum pass: pixels of one slice over 256 threads.
inline int band_blocks(const SliceGeom &g) { return std::max(1, std::min((g.P + 255) / 256, 256)); }

// Groups of sampled slices of the one-call band-maximum pass (grid y): four,
// when there are 32 slices or more.  (More groups make the blocks' global
// atomicMax on the 64 band words contend: 15 groups measured 30 us against
// 22 us for one at 256^3 x 4096, and 8 groups with each block's atomics
// skipped where a larger maximum was already stored 29 us against 16 us for
// four.)
inline int band_groups(const SliceGeom &g, int64_t n_rot) {
  (void)g;
  const int64_t nsl = (n_rot + band_stride(n_rot) - 1) / band_stride(n_rot);
  return nsl >= 32 ? 4 : 1;
}


__global__ void set_word_kernel(unsigned *p, unsigned v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// Bricks some slice plane can reach, for calls with few rotations (a tilt
// series leaves whole wedges of the volume untouched): the brick kernel's
// plane test with twice its margin, so that every brick whose units would
// list a crossing rotation is marked.  Uncrossed bricks get crossed = 0 and
// touched = 0 (no halo shell).
template <int BZ, int BY, int BX>
__global__ __launch_bounds__(256) void brick_cross_kernel(SliceGeom g, BrickGrid bg,
                                                          const PlaneTest *__restrict__ ptest, int n_rot,
                                                          int *__restrict__ crossed, int *__restrict__ touched) {
  const int brick = blockIdx.x * 256 + threadIdx.x;
  if (brick >= bg.nbz * bg.nby * bg.nbx) return;
  const BrickBox bb = brick_box<BZ, BY, BX>(g, bg, brick);
  const double beta[3] = {g.bz, g.by, 0.0};
  float cf[3], hf[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double L = (double)bb.lo[d] - kBoxMargin, U = (double)(bb.lo[d] + bb.ext[d] + 1) + kBoxMargin;
    cf[d] = (float)(0.5 * (L + U) - beta[d]);
    hf[d] = (float)(0.5 * (U - L));
  }
  const float tmarg = 2.0f * (1e-3f + 4e-6f * (fabsf(cf[0]) + fabsf(cf[1]) + fabsf(cf[2]) + hf[0] + hf[1] + hf[2])) +
                      1e-3f;
  int hit = 0;
  for (int r = 0; r < n_rot && !hit; ++r) {
    const float4 q = ptest[r].n;
    const float dist = q.x * cf[0] + q.y * cf[1] + q.z * cf[2];
    const float rad = fabsf(q.x) * hf[0] + fabsf(q.y) * hf[1] + fabsf(q.z) * hf[2];
    hit = (q.w < 0.0f) | (fabsf(dist) <= rad + tmarg);
  }
  crossed[brick] = hit;
  if (!hit) touched[brick] = 0;
}

// Zeros for the voxels of uncrossed bricks, written in every pass instead of
// their units' empty tiles, before the halo pass adds the shells of crossed
// neighbours; an insertion's x = 0 voxels go to its mirror plane.  A voxel v
// on an axis of bricks of B cells belongs to brick min(v / B, nb - 1).
template <int BZ, int BY, int BX>
__global__ __launch_bounds__(256) void uncrossed_fill_kernel(SliceGeom g, BrickGrid bg,
                                                             const int *__restrict__ crossed, ScaleArgs sa,
                                                             f2v *__restrict__ out, float *__restrict__ wout,
                                                             f2v *__restrict__ plane, float *__restrict__ wplane) {
  if (sa.guard && *sa.flag == 0u) return;  // exact redo not needed
  // one block row (z, y) at a time, its x run across the threads
  const int nzy = g.vD * g.vH;
  for (int zy = blockIdx.x; zy < nzy; zy += gridDim.x) {
    const int z = zy / g.vH, y = zy - z * g.vH;
    const int brow = (min(z / BZ, bg.nbz - 1) * bg.nby + min(y / BY, bg.nby - 1)) * bg.nbx;
    const int64_t e0 = (int64_t)zy * g.vW;
    for (int x = threadIdx.x; x < g.vW; x += 256) {
      if (crossed[brow + min(x / BX, bg.nbx - 1)]) continue;
      if (plane && x == 0) {
        plane[zy] = f2v{0.0f, 0.0f};
        wplane[zy] = 0.0f;
      } else {
        out[e0 + x] = f2v{0.0f, 0.0f};
        if (wout) wout[e0 + x] = 0.0f;
      }
    }
  }
}


// x=0-plane Friedel mirror of the insertion, from the plane of direct splats
// (slice_insertion/_insert_central_slices_rfft_3d.py:350-404):
//   V[z,y,0] = P[z,y] + conj(P[mz,my])  unless (mz,my) is DC.
__device__ __forceinline__ void mirror_one(const f2v *__restrict__ plane, const float *__restrict__ wplane,
                                           int D, int H, int Wr, f2v *__restrict__ out,
                                           float *__restrict__ wout, int accumulate, int idx) {
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
    if (wout) w += wout[e];
  }
  out[e] = val;
  if (wout) wout[e] = w;
}

__global__ __launch_bounds__(256) void mirror_kernel(const f2v *__restrict__ plane,
                                                     const float *__restrict__ wplane, int D, int H,
                                                     int Wr, f2v *__restrict__ out,
                                                     float *__restrict__ wout, int accumulate) {
  mirror_one(plane, wplane, D, H, Wr, out, wout, accumulate, (int)(blockIdx.x * 256 + threadIdx.x));
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
  // host copies (the batched plans merge the units of many volumes)
  std::shared_ptr<std::vector<Unit>> h_units;
  std::shared_ptr<std::vector<double>> h_est;
  std::shared_ptr<std::vector<SplitRec>> h_splits;
};

std::mutex g_plan_mutex;
std::map<PlanKey, PlanEntry> g_plans;

// Work per unit is about the total estimate / kTargetSlots: two units per
// resident block slot (256 CUs x 2 blocks).  Measured at 256^3 x 4096
// (extract + adjoint): 512 -> 2.021 ms, 1024 -> 2.009, 2048 -> 2.049,
// 4096 -> 2.343 (smaller units re-read their tiles; larger ones unbalance).
constexpr int kTargetSlots = 256 * 2 * 2;

// Units of the bricks, heaviest first.  The work of a brick is estimated for
// uniformly oriented planes: a plane through DC crosses a brick of
// half-diagonal rb at distance d with probability ~ min(1, rb / d) and covers
// ~ (BZ BY BX)^(2/3) of its pixels.  Bricks no slice sample can reach are
// dropped (gather) or kept as empty units that write zeros (scatter).
// Device copies of a plan's unit and split tables (hipMalloc, kept with the
// cached plan).
int upload_plan(PlanEntry &e) {
  hipError_t err = hipSuccess;
  if (e.n_units > 0 && !e.units) {
    err = hipMalloc((void **)&e.units, e.h_units->size() * sizeof(Unit));
    if (err == hipSuccess)
      err = hipMemcpy(e.units, e.h_units->data(), e.h_units->size() * sizeof(Unit), hipMemcpyHostToDevice);
  }
  if (err == hipSuccess && e.n_splits > 0 && !e.splits) {
    err = hipMalloc((void **)&e.splits, e.h_splits->size() * sizeof(SplitRec));
    if (err == hipSuccess)
      err = hipMemcpy(e.splits, e.h_splits->data(), e.h_splits->size() * sizeof(SplitRec), hipMemcpyHostToDevice);
  }
  if (err == hipSuccess) err = hipGetLastError();
  if (err != hipSuccess) {
    set_error(std::string("brick plan upload (") + std::to_string(e.n_units) + " units, " +
              std::to_string(e.n_splits) + " splits): " + hipGetErrorString(err));
    return TT_ERR_HIP;
  }
  return TT_OK;
}

// device_tables = false: the host tables only (no device allocation; the
// caller copies them where it needs them).
int get_plan(const SliceGeom &g, int BZ, int BY, int BX, bool scatter, int64_t n_rot, PlanEntry &pe,
             bool device_tables = true) {
  int dev = 0;
  if (hipError_t e0 = hipGetDevice(&dev)) {
    if (device_tables) {
      set_error(std::string("brick plan: hipGetDevice: ") + hipGetErrorString(e0));
      return TT_ERR_HIP;
    }
    (void)hipGetLastError();  // host tables only (no device, as in tt_plan_check): clear the query's error
    dev = -1;
  }
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
    if (device_tables)
      if (int st = upload_plan(it->second)) return st;
    pe = it->second;
    if (hipError_t e1 = device_tables ? hipGetLastError() : hipSuccess) {
      set_error(std::string("brick plan (cached): ") + hipGetErrorString(e1));
      return TT_ERR_HIP;
    }
    return TT_OK;
  }
  // bricks tile the cells 0..N-2 of each axis
  const int nbz = (g.vD - 1 + BZ - 1) / BZ, nby = (g.vH - 1 + BY - 1) / BY,
            nbx = (g.vW - 1 + BX - 1) / BX;
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
          const double clo = lo;
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
  e.h_units = std::make_shared<std::vector<Unit>>(host);
  e.h_est = std::make_shared<std::vector<double>>(us.size());
  for (size_t k = 0; k < us.size(); ++k) (*e.h_est)[k] = us[k].est;
  e.h_splits = std::make_shared<std::vector<SplitRec>>(splits);
  if (device_tables)
    if (int st = upload_plan(e)) return st;
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
    if (p) (void)hipFreeAsync(p, s);
  }
};

constexpr int kGBZ = 8, kGBY = 16, kGBX = 16;  // gather brick (cells)
constexpr int kSparseRot = 1024;  // at most this many rotations: bricks no plane reaches are skipped
// adjoint brick (cells): at 256^3 x 4096, 16x8x16 0.955 ms against 0.978 for
// 16x16x8 (8x16x16 was slower; 8x8x32 leaves one block per CU)
constexpr int kABZ = 16, kABY = 8, kABX = 16;
// insertion brick (cells): 24 bytes of accumulators per tile voxel leave room
// for 9 x 9 x 17 voxels; at 256^3 x 4096, 8x8x16 1.43 ms, 16x8x8 1.49,
// 4x16x16 1.56, 4x8x32 1.57 (x-wide bricks also halve the halo pass's lines)
constexpr int kIBZ = 8, kIBY = 8, kIBX = 16;

BrickGrid grid_of(const SliceGeom &g, int BZ, int BY, int BX) {
  BrickGrid bg;
  bg.nbz = (g.vD - 1 + BZ - 1) / BZ;
  bg.nby = (g.vH - 1 + BY - 1) / BY;
  bg.nbx = (g.vW - 1 + BX - 1) / BX;
  return bg;
}

int check_engine(const SliceGeom &g, int64_t n_rot) {
  if (n_rot >= ((int64_t)1 << 30)) return fail_invalid("too many rotations for one call (< 2^30)");
  if (g.H >= 32768) return fail_invalid("slice height must be < 32768");
  if (g.Wr >= 65536) return fail_invalid("slice rfft width must be < 65536");
  return TT_OK;
}


// Scatter driver shared by the adjoint and the insertion.
template <int BZ, int BY, int BX, int MODE>
int run_scatter(const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                float *wout, int accumulate, hipStream_t s) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  // few pixels per brick (a tilt series into a large volume): the binned scatter
  if (binned_selected(g, n_rot)) {
    const int st = binned_scatter(MODE == kInsert, g, src, rot, n_rot, out, wout, accumulate, s);
    if (st != kBinnedNoMemory) return st;
  }
  PlanEntry pe;
  if (int st = get_plan(g, BZ, BY, BX, true, n_rot, pe)) return st;
  const BrickGrid bg = grid_of(g, BZ, BY, BX);
  const int64_t n_bricks = (int64_t)bg.nbz * bg.nby * bg.nbx;
  const size_t np = (size_t)g.vD * g.vH;
  // scratch: words (max slots, flag, band maxima) | plane | wplane |
  // PlaneRec | PlaneTest | rotation-segment slots | halo shells
  size_t off = (kScaleWords * sizeof(unsigned) + 255) & ~(size_t)255;
  const size_t off_plane = off;
  if (MODE == kInsert) off += np * sizeof(f2v);
  const size_t off_wplane = off;
  if (MODE == kInsert) off += np * sizeof(float);
  const size_t zero_end = off;
  off = (off + 255) & ~(size_t)255;
  const size_t off_pre = off;
  off += (size_t)std::max<int64_t>(n_rot, 1) * sizeof(PlaneRec);
  off = (off + 255) & ~(size_t)255;
  const size_t off_test = off;
  off += (size_t)std::max<int64_t>(n_rot, 1) * sizeof(PlaneTest);
  off = (off + 255) & ~(size_t)255;
  const size_t off_seg = off;
  off += (size_t)pe.n_slots * C::NACC * C::NT * sizeof(u64);
  off = (off + 255) & ~(size_t)255;
  const size_t off_halo = off;
  off += (size_t)n_bricks * C::NACC * C::NSH * sizeof(float);
  off = (off + 255) & ~(size_t)255;
  const size_t off_touched = off;  // per brick: reached by a slice (written by every brick)
  off += (size_t)n_bricks * sizeof(int);
  off = (off + 255) & ~(size_t)255;
  // few rotations (a tilt series): test every brick against the planes first
  const bool sparse = n_rot > 0 && n_rot <= kSparseRot;
  const size_t off_crossed = off;
  if (sparse) off += (size_t)n_bricks * sizeof(int);
  Scratch sc(s);
  if (int st = sc.alloc(off)) return st;
  char *base = (char *)sc.p;
  unsigned *words = (unsigned *)base;
  f2v *plane = MODE == kInsert ? (f2v *)(base + off_plane) : nullptr;
  float *wplane = MODE == kInsert ? (float *)(base + off_wplane) : nullptr;
  PlaneRec *pre = (PlaneRec *)(base + off_pre);
  PlaneTest *ptest = (PlaneTest *)(base + off_test);
  u64 *segbuf = (u64 *)(base + off_seg);
  float *halo = (float *)(base + off_halo);
  int *touched = (int *)(base + off_touched);
  int *crossed = sparse ? (int *)(base + off_crossed) : nullptr;
  if (int st = fill_zero<unsigned>(words, (int64_t)(zero_end / sizeof(unsigned)), s)) return st;
  // values the exact maximum reads: the whole (strided) stack, a bound for
  // every channel of a multichannel one
  const int64_t n = (n_rot - 1) * g.ostride + g.P;
  // accumulate: the exact maximum of the input for every brick (the outputs
  // cannot be rewritten); otherwise per-brick scales from sampled band maxima
  // with 8 bits of headroom, and an exact redo if a value exceeded them
  const bool sampled = !accumulate;
  const int mgrid = (int)std::min<int64_t>(std::max<int64_t>(n / 4096, 1), 2048);
  if (n_rot > 0) {
    if (sampled) {
      hipLaunchKernelGGL(band_max_kernel, dim3((unsigned)band_blocks(g), (unsigned)band_groups(g, n_rot)), dim3(256),
                         0, s, src, g, n_rot,
                         band_stride(n_rot), words + kBandWord);
    } else {
      hipLaunchKernelGGL(maxabs_chunks_kernel, dim3(mgrid), dim3(256), 0, s, src, n, (int64_t)1,
                         (const unsigned *)nullptr, words + kSetWords);
    }
    hipLaunchKernelGGL(plane_rec_kernel, dim3((unsigned)((n_rot + 255) / 256)), dim3(256), 0, s, g,
                       rot, (int)n_rot, pre, ptest, words + kSliceWord);
    if (int st = check_launch("band_max / maxabs_chunks / plane_rec")) return st;
  }
  if (sparse) {
    hipLaunchKernelGGL((brick_cross_kernel<BZ, BY, BX>), dim3((unsigned)((n_bricks + 255) / 256)), dim3(256), 0, s,
                       g, bg, (const PlaneTest *)ptest, (int)n_rot, crossed, touched);
    if (int st = check_launch("brick_cross_kernel")) return st;
  }
  for (int pass = 0; pass < (sampled ? 2 : 1); ++pass) {
    ScaleArgs sa;
    sa.maxbits = words + kSetWords;
    sa.bandmax = (sampled && pass == 0) ? words + kBandWord : nullptr;
    sa.flag = words + kFlagWord;
    sa.headroom = (sampled && pass == 0) ? 8 : 0;
    sa.guard = sampled ? pass : 0;
    sa.rotbits = ceil_log2(n_rot);
    sa.slicebits = words + kSliceWord;
    sa.crossed = crossed;
    if (pass == 1 && n_rot > 0) {
      hipLaunchKernelGGL(maxabs_chunks_kernel, dim3(std::min(mgrid, 512)), dim3(256), 0, s, src, n, (int64_t)1,
                         (const unsigned *)(words + kFlagWord), words + kSetWords);
      if (int st = check_launch("maxabs_chunks_kernel")) return st;
    }
    if (sparse && !accumulate) {  // (accumulating: uncrossed voxels keep their values)
      hipLaunchKernelGGL((uncrossed_fill_kernel<BZ, BY, BX>),
                         dim3((unsigned)std::min<int64_t>((int64_t)g.vD * g.vH, pass ? 1024 : 65536)), dim3(256), 0, s, g,
                         bg, (const int *)crossed, sa, out, wout, plane, wplane);
      if (int st = check_launch("uncrossed_fill_kernel")) return st;
    }
    if (pe.n_units > 0) {
      const unsigned ugrid = (unsigned)pe.n_units;
      if (g.use_mask)
        hipLaunchKernelGGL((brick_kernel<BZ, BY, BX, MODE, true>), dim3(ugrid), dim3(C::T),
                           0, s, src, g, bg, (const PlaneRec *)pre, (const PlaneTest *)ptest, (const Unit *)pe.units, (int)n_rot,
                           out, wout, plane, wplane, segbuf, halo, sa, accumulate, touched);
      else
        hipLaunchKernelGGL((brick_kernel<BZ, BY, BX, MODE, false>), dim3(ugrid), dim3(C::T),
                           0, s, src, g, bg, (const PlaneRec *)pre, (const PlaneTest *)ptest, (const Unit *)pe.units, (int)n_rot,
                           out, wout, plane, wplane, segbuf, halo, sa, accumulate, touched);
      if (int st = check_launch("brick_kernel")) return st;
    }
    if (pe.n_splits > 0) {
      hipLaunchKernelGGL((split_combine_kernel<BZ, BY, BX, MODE>),
                         dim3((unsigned)pe.n_splits, (unsigned)((C::NT + 255) / 256)), dim3(256), 0, s, g, bg,
                         (const SplitRec *)pe.splits, (const u64 *)segbuf, sa, accumulate,
                         out, wout, plane, wplane, halo, touched);
      if (int st = check_launch("split_combine_kernel")) return st;
    }
    hipLaunchKernelGGL((halo_combine_kernel<BZ, BY, BX, MODE>),
                       dim3((unsigned)(sa.guard ? std::min<int64_t>(n_bricks, 256) : n_bricks),
                            (unsigned)((C::NSH + 255) / 256)), dim3(256), 0, s,
                       g, bg, (const float *)halo, sa, out, wout, plane, wplane, (const int *)touched);
    if (int st = check_launch("halo_combine_kernel")) return st;
  }
  if (MODE == kInsert) {
    hipLaunchKernelGGL(mirror_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, plane, wplane,
                       g.vD, g.vH, g.vW, out, wout, accumulate);
    if (int st = check_launch("mirror_kernel")) return st;
  }
  return TT_OK;
}

// Host-side check of a brick plan against the buffers run_scatter and the
// gather size from it: every unit's brick, segment and partial slot, every
// split record's slots, the segment-buffer and halo-buffer extents the
// kernels index, and that the units of each brick cover its rotation
// segments exactly once.  stats: n_bricks, n_units, n_splits, n_slots,
// segment-buffer bytes, halo-buffer bytes, empty units, largest nseg.
int plan_check(const SliceGeom &g, int BZ, int BY, int BX, int mode, int64_t n_rot, int64_t stats[8]) {
  PlanEntry pe;
  if (int st = get_plan(g, BZ, BY, BX, mode != kGather, n_rot, pe, false)) return st;
  const BrickGrid bg = grid_of(g, BZ, BY, BX);
  const int64_t n_bricks = (int64_t)bg.nbz * bg.nby * bg.nbx;
  const int64_t NT = (int64_t)(BZ + 1) * (BY + 1) * (BX + 1);
  const int64_t NSH = (int64_t)(BY + 1) * (BX + 1) + (int64_t)BZ * (BX + 1) + (int64_t)BZ * BY;
  const int64_t NA = mode == kGather ? 0 : (mode == kInsert ? 3 : 2);
  const int64_t seg_words = (int64_t)pe.n_slots * NA * NT;  // u64 words of the segment buffer
  auto bad = [](const std::string &m) { return fail_invalid("brick plan check: " + m); };
  std::vector<int> nseg_of(n_bricks, -1), slot_of(n_bricks, -2);
  std::vector<std::vector<char>> cover(n_bricks);
  int64_t empty = 0, max_nseg = 0;
  for (int u = 0; u < pe.n_units; ++u) {
    const Unit &un = (*pe.h_units)[u];
    if (un.brick < 0 || un.brick >= n_bricks) return bad("unit " + std::to_string(u) + ": brick out of range");
    const int nseg = un.seg & 0xffff, sgi = un.seg >> 16;
    if (nseg == 0) {  // an empty scatter brick: writes zeros
      if (mode == kGather || sgi != 0 || un.slot != -1) return bad("unit " + std::to_string(u) + ": bad empty unit");
      ++empty;
    } else if (sgi < 0 || sgi >= nseg) {
      return bad("unit " + std::to_string(u) + ": segment out of range");
    }
    max_nseg = std::max<int64_t>(max_nseg, nseg);
    if (nseg_of[un.brick] == -1) {
      nseg_of[un.brick] = nseg;
      slot_of[un.brick] = un.slot;
      cover[un.brick].assign(std::max(nseg, 1), 0);
    } else if (nseg_of[un.brick] != nseg || slot_of[un.brick] != un.slot) {
      return bad("brick " + std::to_string(un.brick) + ": units disagree on segments or slot");
    }
    if (cover[un.brick][sgi]++) return bad("brick " + std::to_string(un.brick) + ": segment listed twice");
    if (un.slot >= 0) {
      if (mode == kGather || nseg < 2) return bad("unit " + std::to_string(u) + ": slot on an unsplit unit");
      // segbuf + (slot + seg) * NA * NT, NA * NT words
      if ((int64_t)un.slot + nseg > pe.n_slots || ((int64_t)un.slot + sgi + 1) * NA * NT > seg_words)
        return bad("unit " + std::to_string(u) + ": partial slot beyond the segment buffer");
    } else if (mode != kGather && nseg > 1) {
      return bad("unit " + std::to_string(u) + ": split scatter unit without a slot");
    }
  }
  for (int64_t b = 0; b < n_bricks; ++b)
    if (nseg_of[b] > 0)
      for (int k = 0; k < nseg_of[b]; ++k)
        if (cover[b][k] != 1) return bad("brick " + std::to_string(b) + ": segment missing");
  std::vector<char> slot_used(std::max(pe.n_slots, 1), 0);
  for (int k = 0; k < pe.n_splits; ++k) {
    const SplitRec &sp = (*pe.h_splits)[k];
    if (sp.brick < 0 || sp.brick >= n_bricks || sp.nseg < 2 || sp.slot < 0 || (int64_t)sp.slot + sp.nseg > pe.n_slots)
      return bad("split " + std::to_string(k) + ": out of range");
    if (nseg_of[sp.brick] != sp.nseg || slot_of[sp.brick] != sp.slot)
      return bad("split " + std::to_string(k) + ": disagrees with its units");
    for (int q = 0; q < sp.nseg; ++q)
      if (slot_used[sp.slot + q]++) return bad("split " + std::to_string(k) + ": slot shared");
  }
  // the halo buffer: brick * NA * NSH + NA * NSH floats for the last brick
  const int64_t halo_words = n_bricks * NA * NSH;
  if (mode != kGather && (n_bricks - 1) * NA * NSH + NA * NSH > halo_words) return bad("halo buffer");
  if (mode != kGather && (std::max<int64_t>(n_rot, 1) > ((int64_t)1 << 30))) return bad("rotations");
  stats[0] = n_bricks;
  stats[1] = pe.n_units;
  stats[2] = pe.n_splits;
  stats[3] = pe.n_slots;
  stats[4] = seg_words * 8;
  stats[5] = halo_words * 4;
  stats[6] = empty;
  stats[7] = max_nseg;
  return TT_OK;
}

// ---------------------------------------------------------------------------
// Batched launches: many volumes (box sizes and rotation counts of their own)
// in one persistent work-queue launch of the brick kernel per tile
// configuration, plus batched helper kernels.  Everything a run needs is
// planned and allocated when the batch is created, so a run only launches
// kernels (and can be captured in a graph).
// ---------------------------------------------------------------------------
struct BatchItemDev {
  SliceGeom g;
  BrickGrid bg;
  const f2v *src;  // volume (extraction) or slices (insertion)
  const float *rot;
  f2v *out;        // slices (extraction) or volume (insertion)
  float *wout;     // weights (insertion)
  PlaneRec *pre;
  PlaneTest *ptest;
  unsigned *words;  // scale slots (sampled, exact) and the violation flag
  f2v *plane;
  float *wplane;
  u64 *segbuf;
  float *halo;
  int n_rot;
  int n_bricks;
  int64_t n_src;  // values in src (insertion)
};

struct BatchUnit {
  int item, brick, seg, slot;
};

__device__ __forceinline__ ScaleArgs batch_scale(const BatchItemDev &it, int pass) {
  ScaleArgs sa;
  sa.maxbits = it.words + kSetWords;
  sa.bandmax = pass ? nullptr : it.words + kBandWord;
  sa.flag = it.words + kFlagWord;
  sa.headroom = pass ? 0 : 8;
  sa.guard = pass;
  sa.rotbits = ceil_log2(it.n_rot);
  sa.slicebits = it.words + kSliceWord;
  sa.crossed = nullptr;
  return sa;
}

// Persistent: every block takes units from the queue until it is empty.
template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__((Cfg<BZ, BY, BX, MODE>::T), 4) void brick_batch_kernel(
    const BatchItemDev *__restrict__ items, const BatchUnit *__restrict__ units, int n_units,
    unsigned *__restrict__ queue, int pass) {
  __shared__ int s_next;
  for (;;) {
    if (threadIdx.x == 0) s_next = (int)atomicAdd(queue, 1u);
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_next);
    __syncthreads();  // s_next is rewritten by the next iteration
    if (u >= n_units) break;
    const BatchUnit bu = units[u];
    const int ix = __builtin_amdgcn_readfirstlane(bu.item);
    const BatchItemDev &it = items[ix];
    const Unit un{__builtin_amdgcn_readfirstlane(bu.brick), __builtin_amdgcn_readfirstlane(bu.seg),
                  __builtin_amdgcn_readfirstlane(bu.slot)};
    const ScaleArgs sa = batch_scale(it, pass);
    brick_body<BZ, BY, BX, MODE, true>(it.src, it.g, it.bg, it.pre, it.ptest, un, it.n_rot, it.out, it.wout,
                                       it.plane, it.wplane, it.segbuf, it.halo, sa, 0, nullptr);
    __syncthreads();  // the tile is reused by the next unit
  }
}

// blocks: (item, block within the item[, blocks of the item])
__global__ __launch_bounds__(256) void gather_zero_batch_kernel(const BatchItemDev *__restrict__ items,
                                                                const int2 *__restrict__ blocks) {
  const int2 b = blocks[blockIdx.x];
  const BatchItemDev &it = items[b.x];
  gather_zero_body(it.g, it.rot, (int64_t)it.n_rot * it.g.H, it.out, (int64_t)b.y);
}

__global__ __launch_bounds__(256) void plane_rec_batch_kernel(const BatchItemDev *__restrict__ items,
                                                              const int2 *__restrict__ blocks) {
  const int2 b = blocks[blockIdx.x];
  const BatchItemDev &it = items[b.x];
  const int r = b.y * 256 + threadIdx.x;
  const int bits = r < it.n_rot ? plane_rec_one(it.g, it.rot, r, it.pre, it.ptest) : 0;
  max_count_bits(bits, it.words ? it.words + kSliceWord : nullptr);
}

__global__ __launch_bounds__(256) void maxabs_batch_kernel(const BatchItemDev *__restrict__ items,
                                                           const int3 *__restrict__ blocks, int pass) {
  const int3 b = blocks[blockIdx.x];
  const BatchItemDev &it = items[b.x];
  if (pass == 0)
    band_max_body(it.src, it.g, it.n_rot, band_stride(it.n_rot), it.words + kBandWord, b.y, b.z);
  else
    maxabs_body(it.src, it.n_src, 1, (const unsigned *)(it.words + kFlagWord), it.words + kSetWords, b.y, b.z);
}

template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__(256) void split_batch_kernel(const BatchItemDev *__restrict__ items,
                                                          const BatchUnit *__restrict__ splits, int pass) {
  const BatchUnit sp = splits[blockIdx.x];  // item, brick, number of segments, first slot
  const BatchItemDev &it = items[sp.item];
  const SplitRec rec{sp.brick, sp.seg, sp.slot};
  split_body<BZ, BY, BX, MODE>(it.g, it.bg, rec, it.segbuf, batch_scale(it, pass), 0, it.out, it.wout,
                               it.plane, it.wplane, it.halo, (int)(blockIdx.y * 256 + threadIdx.x), nullptr);
}

template <int BZ, int BY, int BX, int MODE>
__global__ __launch_bounds__(256) void halo_batch_kernel(const BatchItemDev *__restrict__ items,
                                                         const int2 *__restrict__ bricks, int pass) {
  const int2 b = bricks[blockIdx.x];  // item, brick
  const BatchItemDev &it = items[b.x];
  halo_body<BZ, BY, BX, MODE>(it.g, it.bg, it.halo, batch_scale(it, pass), it.out, it.wout, it.plane,
                              it.wplane, b.y, (int)(blockIdx.y * 256 + threadIdx.x), nullptr);
}

__global__ __launch_bounds__(256) void mirror_batch_kernel(const BatchItemDev *__restrict__ items,
                                                           const int2 *__restrict__ blocks) {
  const int2 b = blocks[blockIdx.x];
  const BatchItemDev &it = items[b.x];
  mirror_one(it.plane, it.wplane, it.g.vD, it.g.vH, it.g.vW, it.out, it.wout, 0,
             (int)(b.y * 256 + threadIdx.x));
}

// Tile configurations the batch (and its autotuner) can use.
struct TileCfg {
  int bz, by, bx;
};
constexpr TileCfg kGatherCfg[2] = {{kGBZ, kGBY, kGBX}, {8, 8, 8}};
constexpr TileCfg kInsertCfg[2] = {{kIBZ, kIBY, kIBX}, {8, 8, 8}};

struct BatchGroup {
  int cfg = 0;
  int n_units = 0, n_splits = 0, n_bricks = 0;
  BatchUnit *units = nullptr;   // device
  BatchUnit *splits = nullptr;  // device
  int2 *bricks = nullptr;       // device (item, brick) for the shell pass
  unsigned *queue = nullptr;    // two work counters (pass 0, pass 1), in the zeroed region
};

struct BatchPlan {
  int kind = 0;
  int device = 0;
  int grid = 0;  // resident blocks of a persistent launch
  std::vector<tt_batch_item> items;
  std::vector<int> cfg;  // tile configuration of each item
  std::vector<BatchGroup> groups;
  BatchItemDev *d_items = nullptr;
  int2 *zero_blocks = nullptr, *rec_blocks = nullptr, *mirror_blocks = nullptr;
  int3 *max0_blocks = nullptr, *max1_blocks = nullptr;
  int n_zero = 0, n_rec = 0, n_mirror = 0, n_max0 = 0, n_max1 = 0;
  void *zeroed = nullptr;  // counters, scale words, x = 0 planes: cleared by every run
  size_t zeroed_bytes = 0;
  void *scratch = nullptr;  // plane records, partial tiles, halo shells
  std::vector<void *> tables;  // every uploaded table, the item table included
  ~BatchPlan() {
    for (void *p : tables) (void)hipFree(p);
    if (zeroed) (void)hipFree(zeroed);
    if (scratch) (void)hipFree(scratch);
  }
};

int pending(const char *stage) {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return TT_OK;
  set_error(std::string("batch plan (") + stage + "): " + hipGetErrorString(e));
  return TT_ERR_HIP;
}

template <typename T>
int upload(BatchPlan &bp, const std::vector<T> &h, T **d) {
  *d = nullptr;
  if (h.empty()) return TT_OK;
  hipError_t e = hipMalloc((void **)d, h.size() * sizeof(T));
  if (e == hipSuccess) e = hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("batch table upload: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  bp.tables.push_back(*d);
  return TT_OK;
}

template <int BZ, int BY, int BX, int MODE>
void batch_sizes(const SliceGeom &g, int n_slots, size_t &seg_bytes, size_t &halo_bytes, int &n_bricks) {
  typedef Cfg<BZ, BY, BX, MODE> C;
  const BrickGrid bg = grid_of(g, BZ, BY, BX);
  n_bricks = bg.nbz * bg.nby * bg.nbx;
  seg_bytes = MODE == kGather ? 0 : (size_t)n_slots * C::NACC * C::NT * sizeof(u64);
  halo_bytes = MODE == kGather ? 0 : (size_t)n_bricks * C::NACC * C::NSH * sizeof(float);
}

int batch_item_sizes(int kind, int cfg, const SliceGeom &g, int n_slots, size_t &seg, size_t &halo,
                     int &nb) {
  if (kind == TT_BATCH_EXTRACT) {
    if (cfg == 0) batch_sizes<kGatherCfg[0].bz, kGatherCfg[0].by, kGatherCfg[0].bx, kGather>(g, n_slots, seg, halo, nb);
    else batch_sizes<kGatherCfg[1].bz, kGatherCfg[1].by, kGatherCfg[1].bx, kGather>(g, n_slots, seg, halo, nb);
  } else {
    if (cfg == 0) batch_sizes<kInsertCfg[0].bz, kInsertCfg[0].by, kInsertCfg[0].bx, kInsert>(g, n_slots, seg, halo, nb);
    else batch_sizes<kInsertCfg[1].bz, kInsertCfg[1].by, kInsertCfg[1].bx, kInsert>(g, n_slots, seg, halo, nb);
  }
  return TT_OK;
}

SliceGeom batch_geom(const tt_batch_item &it) {
  return make_geom(it.img_d, it.img_h, it.img_w, it.vol_d, it.vol_h, it.vol_wr, it.fftfreq_max, it.x_tolerance);
}

// Builds the plan for the items with the given tile configurations.
int batch_build(BatchPlan &bp) {
  const int n = (int)bp.items.size();
  const bool ext = bp.kind == TT_BATCH_EXTRACT;
  std::vector<SliceGeom> geo(n);
  std::vector<PlanEntry> pes(n);
  std::vector<size_t> seg_b(n), halo_b(n);
  std::vector<int> nbr(n);
  for (int i = 0; i < n; ++i) {
    geo[i] = batch_geom(bp.items[i]);
    const TileCfg tc = ext ? kGatherCfg[bp.cfg[i]] : kInsertCfg[bp.cfg[i]];
    if (int st = get_plan(geo[i], tc.bz, tc.by, tc.bx, !ext, bp.items[i].n_rot, pes[i])) return st;
    if (int st = pending("unit plan")) return st;
    batch_item_sizes(bp.kind, bp.cfg[i], geo[i], pes[i].n_slots, seg_b[i], halo_b[i], nbr[i]);
  }
  // zeroed region: 2 queue words per group, then per item the scale words and
  // (insertion) the x = 0 planes
  const int n_groups = 2;
  size_t zoff = 256;
  std::vector<size_t> words_off(n), plane_off(n), wplane_off(n);
  for (int i = 0; i < n; ++i) {
    words_off[i] = zoff;
    zoff += (kScaleWords * sizeof(unsigned) + 255) & ~(size_t)255;
    const size_t np = (size_t)geo[i].vD * geo[i].vH;
    plane_off[i] = zoff;
    if (!ext) zoff += (np * sizeof(f2v) + 255) & ~(size_t)255;
    wplane_off[i] = zoff;
    if (!ext) zoff += (np * sizeof(float) + 255) & ~(size_t)255;
  }
  // scratch: per item plane records and tests, partial tiles, halo shells
  size_t soff = 0;
  std::vector<size_t> pre_off(n), test_off(n), seg_off(n), halo_off(n);
  for (int i = 0; i < n; ++i) {
    const size_t nr = (size_t)std::max<int64_t>(bp.items[i].n_rot, 1);
    pre_off[i] = soff;
    soff += (nr * sizeof(PlaneRec) + 255) & ~(size_t)255;
    test_off[i] = soff;
    soff += (nr * sizeof(PlaneTest) + 255) & ~(size_t)255;
    seg_off[i] = soff;
    soff += (seg_b[i] + 255) & ~(size_t)255;
    halo_off[i] = soff;
    soff += (halo_b[i] + 255) & ~(size_t)255;
  }
  hipError_t e = hipMalloc(&bp.zeroed, zoff);
  if (e == hipSuccess) e = hipMalloc(&bp.scratch, std::max<size_t>(soff, 256));
  if (e != hipSuccess) {
    set_error(std::string("batch scratch: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  bp.zeroed_bytes = zoff;
  if (int st = pending("scratch")) return st;
  char *z = (char *)bp.zeroed, *sc = (char *)bp.scratch;
  std::vector<BatchItemDev> di(n);
  for (int i = 0; i < n; ++i) {
    const tt_batch_item &it = bp.items[i];
    const TileCfg tc = ext ? kGatherCfg[bp.cfg[i]] : kInsertCfg[bp.cfg[i]];
    BatchItemDev &d = di[i];
    d.g = geo[i];
    d.bg = grid_of(geo[i], tc.bz, tc.by, tc.bx);
    d.src = (const f2v *)it.input;
    d.rot = it.rot_zyx;
    d.out = (f2v *)it.output;
    d.wout = it.weights;
    d.pre = (PlaneRec *)(sc + pre_off[i]);
    d.ptest = (PlaneTest *)(sc + test_off[i]);
    d.words = (unsigned *)(z + words_off[i]);
    d.plane = ext ? nullptr : (f2v *)(z + plane_off[i]);
    d.wplane = ext ? nullptr : (float *)(z + wplane_off[i]);
    d.segbuf = (u64 *)(sc + seg_off[i]);
    d.halo = (float *)(sc + halo_off[i]);
    d.n_rot = (int)it.n_rot;
    d.n_bricks = nbr[i];
    d.n_src = it.n_rot * (int64_t)geo[i].P;
  }
  if (int st = upload(bp, di, &bp.d_items)) return st;
  if (int st = pending("item table")) return st;
  // groups: the units of all items of a configuration, heaviest first
  bp.groups.clear();
  for (int c = 0; c < n_groups; ++c) {
    struct W {
      BatchUnit u;
      double est;
    };
    std::vector<W> w;
    std::vector<BatchUnit> splits;
    std::vector<int2> bricks;
    for (int i = 0; i < n; ++i) {
      if (bp.cfg[i] != c || (ext && bp.items[i].n_rot == 0)) continue;
      const PlanEntry &pe = pes[i];
      for (size_t k = 0; k < pe.h_units->size(); ++k) {
        const Unit &u = (*pe.h_units)[k];
        w.push_back({BatchUnit{i, u.brick, u.seg, u.slot}, (*pe.h_est)[k]});
      }
      for (const SplitRec &sp : *pe.h_splits) splits.push_back(BatchUnit{i, sp.brick, sp.nseg, sp.slot});
      if (!ext)
        for (int b = 0; b < nbr[i]; ++b) bricks.push_back(make_int2(i, b));
    }
    if (w.empty()) continue;
    std::stable_sort(w.begin(), w.end(), [](const W &a, const W &b) { return a.est > b.est; });
    std::vector<BatchUnit> units(w.size());
    for (size_t k = 0; k < w.size(); ++k) units[k] = w[k].u;
    BatchGroup grp;
    grp.cfg = c;
    grp.n_units = (int)units.size();
    grp.n_splits = (int)splits.size();
    grp.n_bricks = (int)bricks.size();
    grp.queue = (unsigned *)z + 2 * c;
    if (int st = upload(bp, units, &grp.units)) return st;
    if (int st = upload(bp, splits, &grp.splits)) return st;
    if (int st = upload(bp, bricks, &grp.bricks)) return st;
    bp.groups.push_back(grp);
    if (int st = pending("unit tables")) return st;
  }
  // helper block tables
  std::vector<int2> zb, rb, mb;
  std::vector<int3> m0, m1;
  for (int i = 0; i < n; ++i) {
    const int64_t nr = bp.items[i].n_rot;
    if (nr == 0) continue;
    for (int64_t b = 0; b < (nr + 255) / 256; ++b) rb.push_back(make_int2(i, (int)b));
    if (ext) {
      const int64_t nb = (nr * geo[i].H + 255) / 256;
      for (int64_t b = 0; b < nb; ++b) zb.push_back(make_int2(i, (int)b));
    } else {
      const int64_t nv = nr * (int64_t)geo[i].P;
      const int s0 = band_blocks(geo[i]);
      const int s1 = (int)std::min<int64_t>(std::max<int64_t>(nv / 4096, 1), 2048);
      for (int b = 0; b < s0; ++b) m0.push_back(make_int3(i, b, s0));
      for (int b = 0; b < s1; ++b) m1.push_back(make_int3(i, b, s1));
    }
  }
  // the insertion's x = 0 mirror pass writes that plane of every volume
  if (!ext)
    for (int i = 0; i < n; ++i) {
      const int64_t np = (int64_t)geo[i].vD * geo[i].vH;
      for (int64_t b = 0; b < (np + 255) / 256; ++b) mb.push_back(make_int2(i, (int)b));
    }
  bp.n_zero = (int)zb.size();
  bp.n_rec = (int)rb.size();
  bp.n_mirror = (int)mb.size();
  bp.n_max0 = (int)m0.size();
  bp.n_max1 = (int)m1.size();
  if (int st = upload(bp, zb, &bp.zero_blocks)) return st;
  if (int st = upload(bp, rb, &bp.rec_blocks)) return st;
  if (int st = upload(bp, mb, &bp.mirror_blocks)) return st;
  if (int st = upload(bp, m0, &bp.max0_blocks)) return st;
  if (int st = upload(bp, m1, &bp.max1_blocks)) return st;
  return pending("block tables");
}

template <int BZ, int BY, int BX>
int batch_run_gather(const BatchPlan &bp, const BatchGroup &grp, hipStream_t s) {
  typedef Cfg<BZ, BY, BX, kGather> C;
  hipLaunchKernelGGL((brick_batch_kernel<BZ, BY, BX, kGather>), dim3((unsigned)std::min(grp.n_units, bp.grid)),
                     dim3(C::T), 0, s, (const BatchItemDev *)bp.d_items, (const BatchUnit *)grp.units, grp.n_units,
                     grp.queue, 0);
  return check_launch("brick_batch_kernel");
}

template <int BZ, int BY, int BX>
int batch_run_insert_pass(const BatchPlan &bp, const BatchGroup &grp, int pass, hipStream_t s) {
  typedef Cfg<BZ, BY, BX, kInsert> C;
  hipLaunchKernelGGL((brick_batch_kernel<BZ, BY, BX, kInsert>), dim3((unsigned)std::min(grp.n_units, bp.grid)),
                     dim3(C::T), 0, s, (const BatchItemDev *)bp.d_items, (const BatchUnit *)grp.units, grp.n_units,
                     grp.queue + pass, pass);
  if (int st = check_launch("brick_batch_kernel")) return st;
  if (grp.n_splits > 0) {
    hipLaunchKernelGGL((split_batch_kernel<BZ, BY, BX, kInsert>),
                       dim3((unsigned)grp.n_splits, (unsigned)((C::NT + 255) / 256)), dim3(256), 0, s,
                       (const BatchItemDev *)bp.d_items, (const BatchUnit *)grp.splits, pass);
    if (int st = check_launch("split_batch_kernel")) return st;
  }
  if (grp.n_bricks > 0) {
    hipLaunchKernelGGL((halo_batch_kernel<BZ, BY, BX, kInsert>),
                       dim3((unsigned)grp.n_bricks, (unsigned)((C::NSH + 255) / 256)), dim3(256), 0, s,
                       (const BatchItemDev *)bp.d_items, (const int2 *)grp.bricks, pass);
    if (int st = check_launch("halo_batch_kernel")) return st;
  }
  return TT_OK;
}

int batch_run(const BatchPlan &bp, hipStream_t s) {
  if (int st = fill_zero<unsigned>((unsigned *)bp.zeroed, (int64_t)(bp.zeroed_bytes / sizeof(unsigned)), s))
    return st;
  if (bp.kind == TT_BATCH_EXTRACT) {
    if (bp.n_zero > 0) {
      hipLaunchKernelGGL(gather_zero_batch_kernel, dim3((unsigned)bp.n_zero), dim3(256), 0, s,
                         (const BatchItemDev *)bp.d_items, (const int2 *)bp.zero_blocks);
      if (int st = check_launch("gather_zero_batch_kernel")) return st;
    }
    if (bp.n_rec > 0) {
      hipLaunchKernelGGL(plane_rec_batch_kernel, dim3((unsigned)bp.n_rec), dim3(256), 0, s,
                         (const BatchItemDev *)bp.d_items, (const int2 *)bp.rec_blocks);
      if (int st = check_launch("plane_rec_batch_kernel")) return st;
    }
    for (const BatchGroup &grp : bp.groups) {
      int st = grp.cfg == 0 ? batch_run_gather<kGatherCfg[0].bz, kGatherCfg[0].by, kGatherCfg[0].bx>(bp, grp, s)
                            : batch_run_gather<kGatherCfg[1].bz, kGatherCfg[1].by, kGatherCfg[1].bx>(bp, grp, s);
      if (st) return st;
    }
    return TT_OK;
  }
  if (bp.n_max0 > 0) {
    hipLaunchKernelGGL(maxabs_batch_kernel, dim3((unsigned)bp.n_max0), dim3(256), 0, s,
                       (const BatchItemDev *)bp.d_items, (const int3 *)bp.max0_blocks, 0);
    if (int st = check_launch("maxabs_batch_kernel")) return st;
  }
  if (bp.n_rec > 0) {
    hipLaunchKernelGGL(plane_rec_batch_kernel, dim3((unsigned)bp.n_rec), dim3(256), 0, s,
                       (const BatchItemDev *)bp.d_items, (const int2 *)bp.rec_blocks);
    if (int st = check_launch("plane_rec_batch_kernel")) return st;
  }
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1 && bp.n_max1 > 0) {
      hipLaunchKernelGGL(maxabs_batch_kernel, dim3((unsigned)bp.n_max1), dim3(256), 0, s,
                         (const BatchItemDev *)bp.d_items, (const int3 *)bp.max1_blocks, 1);
      if (int st = check_launch("maxabs_batch_kernel")) return st;
    }
    for (const BatchGroup &grp : bp.groups) {
      int st = grp.cfg == 0
                   ? batch_run_insert_pass<kInsertCfg[0].bz, kInsertCfg[0].by, kInsertCfg[0].bx>(bp, grp, pass, s)
                   : batch_run_insert_pass<kInsertCfg[1].bz, kInsertCfg[1].by, kInsertCfg[1].bx>(bp, grp, pass, s);
      if (st) return st;
    }
  }
  if (bp.n_mirror > 0) {
    hipLaunchKernelGGL(mirror_batch_kernel, dim3((unsigned)bp.n_mirror), dim3(256), 0, s,
                       (const BatchItemDev *)bp.d_items, (const int2 *)bp.mirror_blocks);
    if (int st = check_launch("mirror_batch_kernel")) return st;
  }
  return TT_OK;
}

int batch_time(BatchPlan &bp, hipStream_t s, float &ms) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    set_error("batch autotune: hipEventCreate failed");
    return TT_ERR_HIP;
  }
  int st = batch_run(bp, s);  // warm-up
  if (!st) {
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < 3 && !st; ++k) st = batch_run(bp, s);
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    (void)hipEventElapsedTime(&ms, e0, e1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

}  // namespace

bool engine_supports(const SliceGeom &g, int64_t n_rot) {
  return n_rot < ((int64_t)1 << 30) && g.H < 32768 && g.Wr < 65536;
}

namespace {

// The gather: plane records into pre/ptest (when the plan has units), then
// one launch of the plan's brick units and the zero pass's row blocks.
template <int NCH>
int gather_launch(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                  PlaneRec *pre, PlaneTest *ptest, const Unit *units, int n_units, hipStream_t s) {
  constexpr int T = Cfg<kGBZ, kGBY, kGBX, kGather>::T;  // one slice row per thread of a zero block
  const int64_t n_zero = (n_rot * g.H + T - 1) / T;
  if (n_zero + n_units >= ((int64_t)1 << 31)) return fail_invalid("too many slice rows for one call");
  if (n_units > 0) {
    hipLaunchKernelGGL(plane_rec_kernel, dim3((unsigned)((n_rot + 255) / 256)), dim3(256), 0, s, g, rot,
                       (int)n_rot, pre, ptest, (unsigned *)nullptr);
    if (int st = check_launch("plane_rec_kernel")) return st;
  }
  const BrickGrid bg = grid_of(g, kGBZ, kGBY, kGBX);
  const dim3 grid((unsigned)(n_units + n_zero));
  if (g.use_mask)
    hipLaunchKernelGGL((gather_kernel<kGBZ, kGBY, kGBX, true, NCH>), grid, dim3(T), 0, s, vol, g, bg,
                       (const PlaneRec *)pre, (const PlaneTest *)ptest, units, n_units, (int)n_rot, rot,
                       (int)n_zero, out);
  else
    hipLaunchKernelGGL((gather_kernel<kGBZ, kGBY, kGBX, false, NCH>), grid, dim3(T), 0, s, vol, g, bg,
                       (const PlaneRec *)pre, (const PlaneTest *)ptest, units, n_units, (int)n_rot, rot,
                       (int)n_zero, out);
  return check_launch("gather_kernel");
}

// Scratch of the allocation-free extraction (engine_extract_noalloc): the
// unit table and the per-rotation plane records, in one file-scope device
// array per GPU (kept until the process ends).  Host code addresses it only
// through the pointer hipGetSymbolAddress returns for the current device.
constexpr int kNoAllocMaxRot = 16384;
constexpr int kNoAllocMaxUnits = 65536;
constexpr size_t kNoAllocPreOff = ((size_t)kNoAllocMaxUnits * sizeof(Unit) + 255) & ~(size_t)255;
constexpr size_t kNoAllocTestOff =
    (kNoAllocPreOff + (size_t)kNoAllocMaxRot * sizeof(PlaneRec) + 255) & ~(size_t)255;
constexpr size_t kNoAllocBytes = kNoAllocTestOff + (size_t)kNoAllocMaxRot * sizeof(PlaneTest);
static __device__ __align__(256) unsigned char noalloc_scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
struct NoAllocDevice {
  char *ptr = nullptr;                      // device address of noalloc_scratch on this GPU
  const std::vector<Unit> *units = nullptr;  // the plan whose units the scratch holds
};
std::mutex g_noalloc_mutex;
NoAllocDevice g_noalloc[kMaxDevices];

}  // namespace

int engine_extract(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                   hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0) return TT_OK;
  PlanEntry pe;
  if (int st = get_plan(g, kGBZ, kGBY, kGBX, false, n_rot, pe)) return st;
  Scratch sc(s);
  PlaneRec *pre = nullptr;
  PlaneTest *ptest = nullptr;
  if (pe.n_units > 0) {
    const size_t off_test = ((size_t)n_rot * sizeof(PlaneRec) + 255) & ~(size_t)255;
    if (int st = sc.alloc(off_test + (size_t)n_rot * sizeof(PlaneTest))) return st;
    pre = (PlaneRec *)sc.p;
    ptest = (PlaneTest *)((char *)sc.p + off_test);
  }
  return gather_launch<1>(g, vol, rot, n_rot, out, pre, ptest, pe.units, pe.n_units, s);
}

int engine_extract_noalloc(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                           hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot > kNoAllocMaxRot)
    return fail_invalid("tt_extract_c64_noalloc: at most " + std::to_string(kNoAllocMaxRot) +
                        " rotations per call (its fixed scratch), got " + std::to_string(n_rot));
  if (n_rot == 0) return TT_OK;
  int dev = 0;
  if (hipError_t e = hipGetDevice(&dev)) {
    set_error(std::string("tt_extract_c64_noalloc: hipGetDevice: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return fail_invalid("tt_extract_c64_noalloc: device index out of range");
  PlanEntry pe;
  if (int st = get_plan(g, kGBZ, kGBY, kGBX, false, n_rot, pe, false)) return st;
  if (pe.n_units > kNoAllocMaxUnits)
    return fail_invalid("tt_extract_c64_noalloc: the volume needs " + std::to_string(pe.n_units) +
                        " brick units, more than its fixed scratch holds (" + std::to_string(kNoAllocMaxUnits) + ")");
  char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    NoAllocDevice &nd = g_noalloc[dev];
    if (!nd.ptr) {  // once per device, with that device current
      void *p = nullptr;
      if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch))) {
        set_error(std::string("tt_extract_c64_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
        return TT_ERR_HIP;
      }
      nd.ptr = (char *)p;
    }
    base = nd.ptr;
    if (pe.n_units > 0 && nd.units != pe.h_units.get()) {
      // stream-ordered: earlier calls on this stream finish with the old units first
      if (hipError_t e = hipMemcpyAsync(base, pe.h_units->data(), pe.h_units->size() * sizeof(Unit),
                                        hipMemcpyHostToDevice, s)) {
        set_error(std::string("tt_extract_c64_noalloc: unit table copy: ") + hipGetErrorString(e));
        return TT_ERR_HIP;
      }
      nd.units = pe.h_units.get();
    }
  }
  return gather_launch<1>(g, vol, rot, n_rot, out, (PlaneRec *)(base + kNoAllocPreOff),
                          (PlaneTest *)(base + kNoAllocTestOff), (const Unit *)base, pe.n_units, s);
}

// Two channels of a multichannel volume (channel stride g.vstride) into a
// multichannel slice stack (rotation stride g.ostride, channel stride g.P):
// one plan, one plane-record pass and one brick launch serve both channels.
int engine_extract2(const SliceGeom &g, const f2v *vol, const float *rot, int64_t n_rot, f2v *out,
                    hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0) return TT_OK;
  PlanEntry pe;
  if (int st = get_plan(g, kGBZ, kGBY, kGBX, false, n_rot, pe)) return st;
  Scratch sc(s);
  PlaneRec *pre = nullptr;
  PlaneTest *ptest = nullptr;
  if (pe.n_units > 0) {
    const size_t off_test = ((size_t)n_rot * sizeof(PlaneRec) + 255) & ~(size_t)255;
    if (int st = sc.alloc(off_test + (size_t)n_rot * sizeof(PlaneTest))) return st;
    pre = (PlaneRec *)sc.p;
    ptest = (PlaneTest *)((char *)sc.p + off_test);
  }
  return gather_launch<2>(g, vol, rot, n_rot, out, pre, ptest, pe.units, pe.n_units, s);
}

int engine_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot, f2v *gV,
                   int accumulate, hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0 && accumulate) return TT_OK;
  return run_scatter<kABZ, kABY, kABX, kAdjoint>(g, gS, rot, n_rot, gV, nullptr, accumulate, s);
}

int engine_plan_check(const SliceGeom &g, int kind, int64_t n_rot, int64_t stats[8]) {
  if (int st = check_engine(g, n_rot)) return st;
  // the batch builder sizes an item's segment and halo buffers itself: they
  // must hold what the plan's kernels index
  auto batch_fits = [&](int bkind, int cfg) -> int {
    size_t seg = 0, halo = 0;
    int nb = 0;
    batch_item_sizes(bkind, cfg, g, (int)stats[3], seg, halo, nb);
    if (nb != stats[0] || seg < (size_t)stats[4] || halo < (size_t)stats[5])
      return fail_invalid("brick plan check: the batch builder's buffers are smaller than the plan needs");
    return TT_OK;
  };
  int st = TT_OK;
  switch (kind) {
    case 0:
      if ((st = plan_check(g, kGBZ, kGBY, kGBX, kGather, n_rot, stats))) return st;
      return batch_fits(TT_BATCH_EXTRACT, 0);
    case 1: return plan_check(g, kABZ, kABY, kABX, kAdjoint, n_rot, stats);
    case 2:
      if ((st = plan_check(g, kIBZ, kIBY, kIBX, kInsert, n_rot, stats))) return st;
      return batch_fits(TT_BATCH_INSERT, 0);
    case 3:
      if ((st = plan_check(g, kGatherCfg[1].bz, kGatherCfg[1].by, kGatherCfg[1].bx, kGather, n_rot, stats))) return st;
      return batch_fits(TT_BATCH_EXTRACT, 1);
    case 4:
      if ((st = plan_check(g, kInsertCfg[1].bz, kInsertCfg[1].by, kInsertCfg[1].bx, kInsert, n_rot, stats))) return st;
      return batch_fits(TT_BATCH_INSERT, 1);
    default:
      return fail_invalid("plan check: kind must be 0 (gather), 1 (adjoint), 2 (insertion), 3 or 4 (the batch's "
                          "alternative gather and insertion tiles)");
  }
}

int engine_insert(const SliceGeom &g, const f2v *slices, const float *rot, int64_t n_rot, f2v *vol,
                  float *w, int accumulate, hipStream_t s) {
  if (int st = check_engine(g, n_rot)) return st;
  if (n_rot == 0 && accumulate) return TT_OK;
  return run_scatter<kIBZ, kIBY, kIBX, kInsert>(g, slices, rot, n_rot, vol, w, accumulate, s);
}

int launch_mirror(const f2v *plane, const float *wplane, const SliceGeom &g, f2v *out, float *wout, int accumulate,
                  hipStream_t s) {
  const size_t np = (size_t)g.vD * g.vH;
  hipLaunchKernelGGL(mirror_kernel, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, plane, wplane, g.vD, g.vH,
                     g.vW, out, wout, accumulate);
  return check_launch("mirror_kernel");
}

}  // namespace tt

// ---------------------------------------------------------------------------
// C ABI of the batched launches
// ---------------------------------------------------------------------------
using namespace tt;

namespace {

int batch_check(int kind, const tt_batch_item &it) {
  if (it.vol_d < 2 || it.vol_h < 2 || it.vol_wr < 2) return fail_invalid("batch item: volume dimensions must be >= 2");
  if (it.img_d < 2 || it.img_h < 2 || it.img_w < 2) return fail_invalid("batch item: image dimensions must be >= 2");
  if (it.n_rot < 0) return fail_invalid("batch item: n_rot must be >= 0");
  if (kind == TT_BATCH_INSERT &&
      (it.img_d != it.vol_d || it.img_h != it.vol_h || it.img_w / 2 + 1 != it.vol_wr))
    return fail_invalid("batch item: insertion needs image shape = the full volume shape");
  if ((int64_t)it.img_h * (it.img_w / 2 + 1) >= (1 << 22)) return fail_invalid("batch item: slice too large");
  if ((int64_t)it.vol_d * it.vol_h * it.vol_wr >= ((int64_t)1 << 31)) return fail_invalid("batch item: volume too large");
  if (it.n_rot > 0 && (!it.rot_zyx || !it.input)) return fail_invalid("batch item: NULL rotations or input");
  // an extraction item without rotations writes nothing (its output may be empty)
  if ((kind == TT_BATCH_INSERT || it.n_rot > 0) && !it.output) return fail_invalid("batch item: NULL output");
  if (kind == TT_BATCH_INSERT && !it.weights) return fail_invalid("batch item: NULL weights");
  const SliceGeom g = batch_geom(it);
  if (!engine_supports(g, it.n_rot)) return fail_invalid("batch item: beyond the brick engine's limits");
  return TT_OK;
}

}  // namespace

extern "C" {

int tt_batch_create(int kind, const tt_batch_item *items, int n_items, int autotune, void *stream,
                    void **handle) {
  if (!handle) return fail_invalid("handle pointer is NULL");
  *handle = nullptr;
  if (kind != TT_BATCH_EXTRACT && kind != TT_BATCH_INSERT) return fail_invalid("kind must be TT_BATCH_EXTRACT or TT_BATCH_INSERT");
  if (n_items <= 0 || !items) return fail_invalid("a batch needs at least one item");
  for (int i = 0; i < n_items; ++i)
    if (int st = batch_check(kind, items[i])) return st;
  hipStream_t s = (hipStream_t)stream;
  int dev = 0, ncu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess || ncu <= 0) {
    set_error(std::string("tt_batch_create: device query: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  if ((e = hipGetLastError()) != hipSuccess) {  // an earlier, unrelated failure: report it, do not run on it
    set_error(std::string("tt_batch_create: pending HIP error: ") + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  auto make = [&](const std::vector<int> &cfg, const std::vector<int> &which, BatchPlan *&out) -> int {
    BatchPlan *bp = new BatchPlan();
    bp->kind = kind;
    bp->device = dev;
    bp->grid = 2 * ncu;
    for (int i : which) bp->items.push_back(items[i]);
    bp->cfg = cfg;
    int st = batch_build(*bp);
    if (st) {
      delete bp;
      return st;
    }
    out = bp;
    return TT_OK;
  };
  std::vector<int> all(n_items), cfg(n_items, 0);
  for (int i = 0; i < n_items; ++i) all[i] = i;
  if (autotune) {
    // per box size (the image shape), time each tile configuration on the
    // items of that size alone and keep the faster
    std::map<std::tuple<int, int, int>, std::vector<int>> sizes;
    for (int i = 0; i < n_items; ++i) sizes[std::make_tuple(items[i].img_d, items[i].img_h, items[i].img_w)].push_back(i);
    for (auto &kv : sizes) {
      float best = 3.0e38f;
      int best_c = 0;
      for (int c = 0; c < 2; ++c) {
        BatchPlan *bp = nullptr;
        if (int st = make(std::vector<int>(kv.second.size(), c), kv.second, bp)) return st;
        float ms = 0.0f;
        if ((e = hipGetLastError()) != hipSuccess) {
          delete bp;
          set_error(std::string("tt_batch_create: after planning: ") + hipGetErrorString(e));
          return TT_ERR_HIP;
        }
        int st = batch_time(*bp, s, ms);
        if (!st && hipStreamSynchronize(s) != hipSuccess) st = check_launch("batch autotune");
        delete bp;
        if (st) return st;
        if (ms < best) {
          best = ms;
          best_c = c;
        }
      }
      for (int i : kv.second) cfg[i] = best_c;
    }
  }
  BatchPlan *bp = nullptr;
  if (int st = make(cfg, all, bp)) return st;
  *handle = bp;
  return TT_OK;
}

int tt_batch_run(void *handle, void *stream) {
  if (!handle) return fail_invalid("batch handle is NULL");
  return batch_run(*(BatchPlan *)handle, (hipStream_t)stream);
}

int tt_batch_tile(void *handle, int item, int *bz, int *by, int *bx) {
  if (!handle || !bz || !by || !bx) return fail_invalid("NULL argument");
  const BatchPlan &bp = *(const BatchPlan *)handle;
  if (item < 0 || item >= (int)bp.items.size()) return fail_invalid("item index out of range");
  const TileCfg tc = bp.kind == TT_BATCH_EXTRACT ? kGatherCfg[bp.cfg[item]] : kInsertCfg[bp.cfg[item]];
  *bz = tc.bz;
  *by = tc.by;
  *bx = tc.bx;
  return TT_OK;
}

int tt_batch_destroy(void *handle) {
  delete (BatchPlan *)handle;
  return TT_OK;
}

}  // extern "C"
