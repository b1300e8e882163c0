// Central-slice gather (extraction) and scatter (insertion) kernels for gfx950,
// with their backward passes, behind the C ABI of include/teamtomo_amd.h.
//
// Reference behaviour (packages/primitives/torch-fourier-slice/src/torch_fourier_slice):
//   extraction  slice_extraction/_extract_central_slices_rfft_3d.py:84-222
//   insertion   slice_insertion/_insert_central_slices_rfft_3d.py:9-106,213-406
// Work mapping: one lane per slice pixel; a block is 256 consecutive pixels of
// one slice, so the rotation is wave-uniform and lives in scalar registers.
// Blocks are remapped so that consecutive slices run on the same XCD and its
// 4 MB L2 sees one slice's footprint of the volume instead of eight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_engine.h"
#include "tt_scatter.h"

namespace tt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail_invalid(const std::string &msg) {
  set_error(msg);
  return TT_ERR_INVALID;
}

int check_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return TT_ERR_HIP;
  }
  return TT_OK;
}

SliceGeom make_geom(int img_d, int img_h, int img_w, int vol_d, int vol_h, int vol_wr,
                    double fftfreq_max, double x_tolerance) {
  SliceGeom g;
  g.H = img_h;
  g.Wr = img_w / 2 + 1;
  g.half_H = img_h / 2;
  g.P = g.H * g.Wr;
  g.inv_H = 1.0 / (double)img_h;
  g.inv_W = 1.0 / (double)img_w;
  g.az = (double)(img_d - 1) / (1.0 - 1.0 / (double)img_d);
  g.ay = (double)(img_h - 1) / (1.0 - 1.0 / (double)img_h);
  g.ax = 2.0 * (double)(g.Wr - 1);
  g.bz = 0.5 * g.az;
  g.by = 0.5 * g.ay;
  g.vD = vol_d;
  g.vH = vol_h;
  g.vW = vol_wr;
  g.use_mask = fftfreq_max >= 0.0 ? 1 : 0;
  g.fmax = (float)fftfreq_max;
  g.inv_H32 = (float)(1.0 / (double)img_h);
  g.inv_W32 = (float)(1.0 / (double)img_w);
  g.xtol = (double)(float)x_tolerance;
  g.xtol10 = (double)(float)(x_tolerance * 10.0);
  g.tolx = g.xtol * g.ax;
  g.invWr_f = 1.0f / (float)g.Wr;
  g.ostride = (int64_t)g.P;
  g.vstride = (int64_t)vol_d * vol_h * vol_wr;
  return g;
}

constexpr int kBlock = 256;

// Logical block id with XCD-aware remapping.  The dispatcher hands block b to
// XCD b % 8; the remap gives XCD k the contiguous logical range
// [k * nb / 8, (k + 1) * nb / 8) when nb is a multiple of 8.
__device__ __forceinline__ int64_t logical_block(int64_t nb_padded) {
  int64_t b = blockIdx.x;
  int64_t per = nb_padded >> 3;
  return (b & 7) * per + (b >> 3);
}

// ---------------------------------------------------------------------------
// Extraction forward: trilinear gather with Friedel conjugation.
// Each lane reads the x-adjacent corner pair with one 16-byte load, so the 8
// corners cost 4 vector memory instructions.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void extract_fwd_kernel(
    const f2v *__restrict__ vol, SliceGeom g, const float *__restrict__ rot, int blocks_per_rot,
    int64_t n_blocks, int64_t nb_padded, f2v *__restrict__ out) {
  int64_t lb = logical_block(nb_padded);
  if (lb >= n_blocks) return;
  int64_t r_idx = lb / blocks_per_rot;
  int p = (int)(lb - r_idx * blocks_per_rot) * kBlock + (int)threadIdx.x;
  if (p >= g.P) return;
  Rot r = load_rot(rot + r_idx * 9, g);
  int i, j;
  split_pixel(g, p, i, j);
  f2v res = {0.0f, 0.0f};
  Coord c = slice_coord<false>(g, r, i, j);
  if (in_disc(g, i, j) && inside(g, c)) {
    Cell k = cell_of(g, c);
    const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
    const f2v *base = vol + k.z0 * sz + k.y0 * sy + k.x0;
    f4a8 a00 = *reinterpret_cast<const f4a8 *>(base);
    f4a8 a01 = *reinterpret_cast<const f4a8 *>(base + sy);
    f4a8 a10 = *reinterpret_cast<const f4a8 *>(base + sz);
    f4a8 a11 = *reinterpret_cast<const f4a8 *>(base + sz + sy);
    res = trilerp(f2v{a00.x, a00.y}, f2v{a00.z, a00.w}, f2v{a01.x, a01.y}, f2v{a01.z, a01.w},
                  f2v{a10.x, a10.y}, f2v{a10.z, a10.w}, f2v{a11.x, a11.y}, f2v{a11.z, a11.w}, k.tz,
                  k.ty, k.tx);
    if (c.conj) res.y = -res.y;
  }
  out[r_idx * g.ostride + p] = res;
}

// Block-wide sum of 6 doubles, lane 0 of the block adds them to gR.
__device__ __forceinline__ void block_add_rot_grad(double acc[6], double *__restrict__ gR) {
  __shared__ double red[kBlock / 64][6];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    acc[q] = v;
  }
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    for (int q = 0; q < 6; ++q) red[wid][q] = acc[q];
  __syncthreads();
  if (threadIdx.x < 6) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += red[w][threadIdx.x];
    // acc layout: (row d, column y|x) -> gR[d*3 + 1 + col]
    int d = threadIdx.x >> 1, col = threadIdx.x & 1;
    unsafeAtomicAdd(gR + d * 3 + 1 + col, s);
  }
}

// ---------------------------------------------------------------------------
// Extraction backward (pixel-driven): adjoint scatter into gV with global
// float atomics and/or the analytic rotation gradient.
// dL/dc_d = Re(conj(gS) kappa(dS/dc_d)), zero where c_d <= 0 or c_d >= N-1
// (grid_sample's border-clip gradient); dc_d/dR[d][col] = a_d s v_col.
// ---------------------------------------------------------------------------
template <bool GV, bool GR>
__global__ __launch_bounds__(kBlock) void extract_bwd_kernel(
    const f2v *__restrict__ gS, const f2v *__restrict__ vol, SliceGeom g,
    const float *__restrict__ rot, int blocks_per_rot, int64_t n_blocks, int64_t nb_padded,
    f2v *__restrict__ gV, double *__restrict__ gR) {
  int64_t lb = logical_block(nb_padded);
  if (lb >= n_blocks) return;
  int64_t r_idx = lb / blocks_per_rot;
  int p = (int)(lb - r_idx * blocks_per_rot) * kBlock + (int)threadIdx.x;
  Rot r = load_rot(rot + r_idx * 9, g);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (p < g.P) {
    int i, j;
    split_pixel(g, p, i, j);
    Coord c = slice_coord<false>(g, r, i, j);
    if (in_disc(g, i, j) && inside(g, c)) {
      Cell k = cell_of(g, c);
      f2v gs = gS[r_idx * g.P + p];
      if (c.conj) gs.y = -gs.y;  // kappa(gS)
      const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
      const int64_t b0 = k.z0 * sz + k.y0 * sy + k.x0;
      float ux = 1.0f - k.tx, uy = 1.0f - k.ty, uz = 1.0f - k.tz;
      if (GV) {
        float *gvf = reinterpret_cast<float *>(gV);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
          float w = (dz ? k.tz : uz) * (dy ? k.ty : uy) * (dx ? k.tx : ux);
          int64_t e = b0 + dz * sz + dy * sy + dx;
          unsafeAtomicAdd(gvf + 2 * e, w * gs.x);
          unsafeAtomicAdd(gvf + 2 * e + 1, w * gs.y);
        }
      }
      if (GR) {
        const f2v *base = vol + b0;
        f4a8 a00 = *reinterpret_cast<const f4a8 *>(base);
        f4a8 a01 = *reinterpret_cast<const f4a8 *>(base + sy);
        f4a8 a10 = *reinterpret_cast<const f4a8 *>(base + sz);
        f4a8 a11 = *reinterpret_cast<const f4a8 *>(base + sz + sy);
        // x-derivative and x-lerp of each (z, y) corner pair
        float dxr00 = a00.z - a00.x, dxi00 = a00.w - a00.y;
        float dxr01 = a01.z - a01.x, dxi01 = a01.w - a01.y;
        float dxr10 = a10.z - a10.x, dxi10 = a10.w - a10.y;
        float dxr11 = a11.z - a11.x, dxi11 = a11.w - a11.y;
        float lr00 = a00.x + k.tx * dxr00, li00 = a00.y + k.tx * dxi00;
        float lr01 = a01.x + k.tx * dxr01, li01 = a01.y + k.tx * dxi01;
        float lr10 = a10.x + k.tx * dxr10, li10 = a10.y + k.tx * dxi10;
        float lr11 = a11.x + k.tx * dxr11, li11 = a11.y + k.tx * dxi11;
        // dS/dz, dS/dy, dS/dx (complex), trilinear with floor/floor+1 corners
        float dzr = uy * (lr10 - lr00) + k.ty * (lr11 - lr01);
        float dzi = uy * (li10 - li00) + k.ty * (li11 - li01);
        float dyr = uz * (lr01 - lr00) + k.tz * (lr11 - lr10);
        float dyi = uz * (li01 - li00) + k.tz * (li11 - li10);
        float dxr = uz * (uy * dxr00 + k.ty * dxr01) + k.tz * (uy * dxr10 + k.ty * dxr11);
        float dxi = uz * (uy * dxi00 + k.ty * dxi01) + k.tz * (uy * dxi10 + k.ty * dxi11);
        // Re(conj(gS) kappa(z)) = Re(conj(kappa(gS)) z) for the conj-folded gs
        float gz = gs.x * dzr + gs.y * dzi;
        float gy = gs.x * dyr + gs.y * dyi;
        float gx = gs.x * dxr + gs.y * dxi;
        if (!(c.z > 0.0 && c.z < (double)(g.vD - 1))) gz = 0.0f;
        if (!(c.y > 0.0 && c.y < (double)(g.vH - 1))) gy = 0.0f;
        if (!(c.x > 0.0 && c.x < (double)(g.vW - 1))) gx = 0.0f;
        double s = c.conj ? -1.0 : 1.0;
        double fy = (double)(i - g.half_H) * g.inv_H;
        double fx = (double)j * g.inv_W;
        double tzv = (double)gz * g.az * s, tyv = (double)gy * g.ay * s, txv = (double)gx * g.ax * s;
        acc[0] = tzv * fy;
        acc[1] = tzv * fx;
        acc[2] = tyv * fy;
        acc[3] = tyv * fx;
        acc[4] = txv * fy;
        acc[5] = txv * fx;
      }
    }
  }
  if (GR) block_add_rot_grad(acc, gR + r_idx * 9);
}

// ---------------------------------------------------------------------------
// Insertion backward: adjoint gather (grad w.r.t. slices) including the
// Friedel mirror, and/or the rotation gradient through the trilinear weights.
// ---------------------------------------------------------------------------
template <typename T, bool GS, bool GR>
__global__ __launch_bounds__(kBlock) void insert_bwd_kernel(
    const typename Cx<T>::type *__restrict__ gV, const T *__restrict__ gW,
    const typename Cx<T>::type *__restrict__ slices, SliceGeom g, const float *__restrict__ rot,
    int blocks_per_rot, int64_t n_blocks, int64_t nb_padded, typename Cx<T>::type *__restrict__ gS,
    double *__restrict__ gR) {
  typedef typename Cx<T>::type C2;
  int64_t lb = logical_block(nb_padded);
  if (lb >= n_blocks) return;
  int64_t r_idx = lb / blocks_per_rot;
  int p = (int)(lb - r_idx * blocks_per_rot) * kBlock + (int)threadIdx.x;
  Rot r = load_rot(rot + r_idx * 9, g);
  double acc[6] = {0, 0, 0, 0, 0, 0};
  if (p < g.P) {
    int i, j;
    split_pixel(g, p, i, j);
    Coord c = slice_coord<true>(g, r, i, j);
    bool ok = !(j == 0 && i < g.half_H) && in_disc(g, i, j) && inside(g, c);
    C2 out = {0, 0};
    if (ok) {
      Cell k = cell_of(g, c);
      const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
      T tz = (T)k.tz, ty = (T)k.ty, tx = (T)k.tx;
      T uz = (T)1 - tz, uy = (T)1 - ty, ux = (T)1 - tx;
      C2 d = {0, 0};
      if (GR) {
        d = slices[r_idx * g.P + p];
        if (c.conj) d.y = -d.y;
      }
      C2 acc_g = {0, 0};
      double qd[3] = {0, 0, 0};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
        int z = k.z0 + dz, y = k.y0 + dy, x = k.x0 + dx;
        T wz = dz ? tz : uz, wy = dy ? ty : uy, wx = dx ? tx : ux;
        T w = wz * wy * wx;
        C2 v = gV[(int64_t)z * sz + (int64_t)y * sy + x];
        acc_g += w * v;
        T qv = 0;
        if (GR) qv = v.x * d.x + v.y * d.y + (gW ? gW[(int64_t)z * sz + (int64_t)y * sy + x] : (T)0);
        if (x == 0 && !(z == g.vD / 2 && y == g.vH / 2)) {
          int mz = (g.vD - z) % g.vD, my = (g.vH - y) % g.vH;
          C2 m = gV[(int64_t)mz * sz + (int64_t)my * sy];
          acc_g.x += w * m.x;
          acc_g.y -= w * m.y;
          // Re(conj(gV[mirror]) conj(d)) = m.x d.x - m.y d.y
          if (GR) qv += m.x * d.x - m.y * d.y + (gW ? gW[(int64_t)mz * sz + (int64_t)my * sy] : (T)0);
        }
        if (GR) {
          T sgz = dz ? (T)1 : (T)-1, sgy = dy ? (T)1 : (T)-1, sgx = dx ? (T)1 : (T)-1;
          qd[0] += (double)(sgz * wy * wx * qv);
          qd[1] += (double)(wz * sgy * wx * qv);
          qd[2] += (double)(wz * wy * sgx * qv);
        }
      }
      out = acc_g;
      if (c.conj) out.y = -out.y;
      if (GR) {
        // floor == ceil at integer coordinates: the two corners coincide and
        // the derivative along that axis vanishes (torch differentiates the
        // weights only, slice_insertion/...:298-329).
        if (c.z == floor(c.z)) qd[0] = 0.0;
        if (c.y == floor(c.y)) qd[1] = 0.0;
        if (c.x == floor(c.x)) qd[2] = 0.0;
        double s = c.conj ? -1.0 : 1.0;
        double fy = (double)(i - g.half_H) * g.inv_H;
        double fx = (double)j * g.inv_W;
        double tzv = qd[0] * g.az * s, tyv = qd[1] * g.ay * s, txv = qd[2] * g.ax * s;
        acc[0] = tzv * fy;
        acc[1] = tzv * fx;
        acc[2] = tyv * fy;
        acc[3] = tyv * fx;
        acc[4] = txv * fy;
        acc[5] = txv * fx;
      }
    }
    if (GS) gS[r_idx * g.P + p] = out;
  }
  if (GR) block_add_rot_grad(acc, gR + r_idx * 9);
}

// ---------------------------------------------------------------------------
// Host launch helpers
// ---------------------------------------------------------------------------
struct Launch {
  int blocks_per_rot;
  int64_t n_blocks, nb_padded;
};

static Launch plan(const SliceGeom &g, int64_t n_rot) {
  Launch l;
  l.blocks_per_rot = (g.P + kBlock - 1) / kBlock;
  l.n_blocks = (int64_t)l.blocks_per_rot * n_rot;
  l.nb_padded = (l.n_blocks + 7) & ~(int64_t)7;
  return l;
}

static int check_common(int vol_d, int vol_h, int vol_wr, int img_d, int img_h, int img_w,
                        int64_t n_rot, const void *rot) {
  if (vol_d < 2 || vol_h < 2 || vol_wr < 2)
    return fail_invalid("volume dimensions must be >= 2 (rfft width >= 2)");
  if (img_d < 2 || img_h < 2 || img_w < 2) return fail_invalid("image_shape dimensions must be >= 2");
  if (n_rot < 0) return fail_invalid("n_rot must be >= 0");
  if (n_rot > 0 && rot == nullptr) return fail_invalid("rotation pointer is NULL");
  int64_t P = (int64_t)img_h * (img_w / 2 + 1);
  if (P >= (1 << 22)) return fail_invalid("slice has too many pixels (H*(W/2+1) must be < 2^22)");
  if ((int64_t)vol_d * vol_h * vol_wr >= ((int64_t)1 << 31))
    return fail_invalid("volume has too many voxels (must be < 2^31)");
  return TT_OK;
}

// Grid size cap: split very large stacks into several launches of whole slices.
static const int64_t kMaxBlocksPerLaunch = (int64_t)1 << 23;

}  // namespace tt

using namespace tt;

extern "C" {

const char *tt_last_error(void) { return g_last_error.c_str(); }

int tt_abi_version(void) { return 13; }

static int g_gather_path = TT_GATHER_AUTO;

int tt_set_gather_path(int path) {
  if (path != TT_GATHER_AUTO && path != TT_GATHER_PIXEL && path != TT_GATHER_BRICK)
    return fail_invalid("gather path must be TT_GATHER_AUTO, TT_GATHER_PIXEL or TT_GATHER_BRICK");
  g_gather_path = path;
  return TT_OK;
}

int tt_set_scatter_path(int path) { return set_scatter_path(path); }

int tt_plan_check(int kind, int vol_d, int vol_h, int vol_wr, int img_d, int img_h, int img_w, int64_t n_rot,
                  double fftfreq_max, int64_t *stats) {
  if (!stats) return fail_invalid("stats pointer is NULL");
  const float dummy = 0.0f;  // no rotations are read
  int st = check_common(vol_d, vol_h, vol_wr, img_d, img_h, img_w, n_rot, &dummy);
  if (st) return st;
  SliceGeom g = make_geom(img_d, img_h, img_w, vol_d, vol_h, vol_wr, fftfreq_max, 1e-8);
  return engine_plan_check(g, kind, n_rot, stats);
}

int tt_extract_c64(const void *volume, int vol_d, int vol_h, int vol_wr, int img_d, int img_h,
                   int img_w, const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                   double x_tolerance, void *slices, void *stream) {
  int st = check_common(vol_d, vol_h, vol_wr, img_d, img_h, img_w, n_rot, rot_zyx);
  if (st) return st;
  if (n_rot == 0) return TT_OK;
  if (!volume || !slices) return fail_invalid("volume / slices pointer is NULL");
  SliceGeom g = make_geom(img_d, img_h, img_w, vol_d, vol_h, vol_wr, fftfreq_max, x_tolerance);
  hipStream_t s = (hipStream_t)stream;
  // the brick engine, unless the pixel gather is asked for or (automatic
  // choice) the geometry is beyond the engine's limits
  if (g_gather_path == TT_GATHER_BRICK || (g_gather_path == TT_GATHER_AUTO && engine_supports(g, n_rot)))
    return engine_extract(g, (const f2v *)volume, rot_zyx, n_rot, (f2v *)slices, s);
  int64_t bpr = (g.P + kBlock - 1) / kBlock;
  int64_t rots_per_launch = std::max<int64_t>(1, kMaxBlocksPerLaunch / bpr);
  for (int64_t r0 = 0; r0 < n_rot; r0 += rots_per_launch) {
    int64_t nr = std::min(rots_per_launch, n_rot - r0);
    Launch l = plan(g, nr);
    hipLaunchKernelGGL(extract_fwd_kernel, dim3((unsigned)l.nb_padded), dim3(kBlock), 0, s,
                       (const f2v *)volume, g, rot_zyx + r0 * 9, l.blocks_per_rot, l.n_blocks,
                       l.nb_padded, (f2v *)slices + r0 * g.P);
    if ((st = check_launch("extract_fwd_kernel"))) return st;
  }
  return TT_OK;
}

// tt_extract_c64 without any allocation: the brick gather with its plane
// records and unit table in a fixed device array (see the header).
int tt_extract_c64_noalloc(const void *volume, int vol_d, int vol_h, int vol_wr, int img_d, int img_h,
                           int img_w, const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                           double x_tolerance, void *slices, void *stream) {
  int st = check_common(vol_d, vol_h, vol_wr, img_d, img_h, img_w, n_rot, rot_zyx);
  if (st) return st;
  if (n_rot == 0) return TT_OK;
  if (!volume || !slices) return fail_invalid("volume / slices pointer is NULL");
  SliceGeom g = make_geom(img_d, img_h, img_w, vol_d, vol_h, vol_wr, fftfreq_max, x_tolerance);
  return engine_extract_noalloc(g, (const f2v *)volume, rot_zyx, n_rot, (f2v *)slices, (hipStream_t)stream);
}

int tt_extract_mc_c64(const void *volume, int channels, int vol_d, int vol_h, int vol_wr, int img_d,
                      int img_h, int img_w, const float *rot_zyx, int64_t n_rot, double fftfreq_max,
                      double x_tolerance, void *slices, void *stream) {
  int st = check_common(vol_d, vol_h, vol_wr, img_d, img_h, img_w, n_rot, rot_zyx);
  if (st) return st;
  if (channels < 1) return fail_invalid("channels must be >= 1");
  if (n_rot == 0) return TT_OK;
  if (!volume || !slices) return fail_invalid("volume / slices pointer is NULL");
  SliceGeom g = make_geom(img_d, img_h, img_w, vol_d, vol_h, vol_wr, fftfreq_max, x_tolerance);
  if (channels == 1)
    return tt_extract_c64(volume, vol_d, vol_h, vol_wr, img_d, img_h, img_w, rot_zyx, n_rot, fftfreq_max,
                          x_tolerance, slices, stream);
  g.ostride = (int64_t)channels * g.P;
  hipStream_t s = (hipStream_t)stream;
  const f2v *vol = (const f2v *)volume;
  f2v *out = (f2v *)slices;
  if (g_gather_path == TT_GATHER_BRICK || (g_gather_path == TT_GATHER_AUTO && engine_supports(g, n_rot))) {
    // channels two at a time through one brick launch each; an odd last
    // channel is taken with its predecessor (which is written again, equal)
    for (int c0 = 0; c0 < channels; c0 += 2) {
      const int cc = std::min(c0, channels - 2);
      if ((st = engine_extract2(g, vol + cc * g.vstride, rot_zyx, n_rot, out + (int64_t)cc * g.P, s))) return st;
    }
    return TT_OK;
  }
  int64_t bpr = (g.P + kBlock - 1) / kBlock;
  int64_t rots_per_launch = std::max<int64_t>(1, kMaxBlocksPerLaunch / bpr);
  for (int c = 0; c < channels; ++c)
    for (int64_t r0 = 0; r0 < n_rot; r0 += rots_per_launch) {
      int64_t nr = std::min(rots_per_launch, n_rot - r0);
      Launch l = plan(g, nr);
      hipLaunchKernelGGL(extract_fwd_kernel, dim3((unsigned)l.nb_padded), dim3(kBlock), 0, s,
                         vol + c * g.vstride, g, rot_zyx + r0 * 9, l.blocks_per_rot, l.n_blocks, l.nb_padded,
                         out + r0 * g.ostride + (int64_t)c * g.P);
      if ((st = check_launch("extract_fwd_kernel"))) return st;
    }
  return TT_OK;
}

int tt_extract_bwd_c64(const void *grad_slices, const void *volume, int vol_d, int vol_h,
                       int vol_wr, int img_d, int img_h, int img_w, const float *rot_zyx,
                       int64_t n_rot, double fftfreq_max, double x_tolerance, void *grad_volume,
                       double *grad_rot, int flags, void *stream) {
  int st = check_common(vol_d, vol_h, vol_wr, img_d, img_h, img_w, n_rot, rot_zyx);
  if (st) return st;
  if (n_rot == 0) {  // no slices: an overwritten volume gradient is zero
    if (grad_volume && (flags & TT_FLAG_OVERWRITE))
      return zero_fill<float>((float *)grad_volume, 2 * (int64_t)vol_d * vol_h * vol_wr, (hipStream_t)stream);
    return TT_OK;
  }
  if (!grad_volume && !grad_rot) return TT_OK;
  if (!grad_slices) return fail_invalid("grad_slices pointer is NULL");
  if (grad_rot && !volume) return fail_invalid("volume is required for the rotation gradient");
  SliceGeom g = make_geom(img_d, img_h, img_w, vol_d, vol_h, vol_wr, fftfreq_max, x_tolerance);
  hipStream_t s = (hipStream_t)stream;
  if (grad_volume) {
    st = scatter_extract_adjoint(g, (const f2v *)grad_slices, rot_zyx, n_rot, (f2v *)grad_volume,
                                 flags, s);
    if (st) return st;
  }
  if (grad_rot) {
    int64_t bpr = (g.P + kBlock - 1) / kBlock;
    int64_t rots_per_launch = std::max<int64_t>(1, kMaxBlocksPerLaunch / bpr);
    for (int64_t r0 = 0; r0 < n_rot; r0 += rots_per_launch) {
      int64_t nr = std::min(rots_per_launch, n_rot - r0);
      Launch l = plan(g, nr);
      hipLaunchKernelGGL((extract_bwd_kernel<false, true>), dim3((unsigned)l.nb_padded),
                         dim3(kBlock), 0, s, (const f2v *)grad_slices + r0 * g.P,
                         (const f2v *)volume, g, rot_zyx + r0 * 9, l.blocks_per_rot, l.n_blocks,
                         l.nb_padded, (f2v *)nullptr, grad_rot + r0 * 9);
      if ((st = check_launch("extract_bwd_kernel"))) return st;
    }
  }
  return TT_OK;
}

}  // extern "C"

template <typename T>
static int insert_impl(const void *slices, int vol_d, int vol_h, int vol_w, const float *rot_zyx,
                       int64_t n_rot, double fftfreq_max, void *volume, T *weights,
                       int flags, void *stream) {
  int vol_wr = vol_w / 2 + 1;
  int st = check_common(vol_d, vol_h, vol_wr, vol_d, vol_h, vol_w, n_rot, rot_zyx);
  if (st) return st;
  if (n_rot == 0) {  // no slices: overwritten outputs are zero
    if (!(flags & TT_FLAG_OVERWRITE)) return TT_OK;
    if (!volume || !weights) return fail_invalid("volume / weights pointer is NULL");
    const int64_t nv = (int64_t)vol_d * vol_h * vol_wr;
    if ((st = zero_fill<T>((T *)volume, 2 * nv, (hipStream_t)stream))) return st;
    return zero_fill<T>(weights, nv, (hipStream_t)stream);
  }
  if (!slices || !volume || !weights) return fail_invalid("slices / volume / weights pointer is NULL");
  SliceGeom g = make_geom(vol_d, vol_h, vol_w, vol_d, vol_h, vol_wr, fftfreq_max, 1e-8);
  return scatter_insert<T>(g, (const typename Cx<T>::type *)slices, rot_zyx, n_rot,
                           (typename Cx<T>::type *)volume, weights, flags,
                           (hipStream_t)stream);
}

extern "C" {

int tt_insert_c64(const void *slices, int vol_d, int vol_h, int vol_w, const float *rot_zyx,
                  int64_t n_rot, double fftfreq_max, void *volume, float *weights,
                  int flags, void *stream) {
  return insert_impl<float>(slices, vol_d, vol_h, vol_w, rot_zyx, n_rot, fftfreq_max, volume,
                            weights, flags, stream);
}

// Multichannel insertion: every channel through the brick engine at the same
// coordinates, reading its channel of the (n, C, H, Wr) stack in place; the
// weights (which do not depend on the values) are accumulated with channel 0
// only.
int tt_insert_mc_c64(const void *slices, int channels, int vol_d, int vol_h, int vol_w, const float *rot_zyx,
                     int64_t n_rot, double fftfreq_max, void *volume, float *weights, int flags, void *stream) {
  const int vol_wr = vol_w / 2 + 1;
  int st = check_common(vol_d, vol_h, vol_wr, vol_d, vol_h, vol_w, n_rot, rot_zyx);
  if (st) return st;
  if (channels < 1) return fail_invalid("channels must be >= 1");
  if (!volume || !weights) return fail_invalid("volume / weights pointer is NULL");
  if (n_rot > 0 && !slices) return fail_invalid("slices pointer is NULL");
  if (flags & TT_FLAG_PIXEL_ATOMIC) return fail_invalid("tt_insert_mc_c64 runs the brick engine only");
  SliceGeom g = make_geom(vol_d, vol_h, vol_w, vol_d, vol_h, vol_wr, fftfreq_max, 1e-8);
  if (!engine_supports(g, n_rot)) return fail_invalid("geometry beyond the brick engine's limits");
  hipStream_t s = (hipStream_t)stream;
  const bool overwrite = flags & TT_FLAG_OVERWRITE;
  const int64_t nv = (int64_t)vol_d * vol_h * vol_wr;
  if (n_rot == 0) {
    if (!overwrite) return TT_OK;
    if ((st = zero_fill<float>((float *)volume, 2 * nv * channels, s))) return st;
    return zero_fill<float>(weights, nv, s);
  }
  g.ostride = (int64_t)channels * g.P;
  for (int c = 0; c < channels; ++c) {
    st = engine_insert(g, (const f2v *)slices + (int64_t)c * g.P, rot_zyx, n_rot, (f2v *)volume + c * nv,
                       c == 0 ? weights : nullptr, !overwrite, s);
    if (st) return st;
  }
  return TT_OK;
}

int tt_insert_c128(const void *slices, int vol_d, int vol_h, int vol_w, const float *rot_zyx,
                   int64_t n_rot, double fftfreq_max, void *volume, double *weights,
                   int flags, void *stream) {
  return insert_impl<double>(slices, vol_d, vol_h, vol_w, rot_zyx, n_rot, fftfreq_max, volume,
                             weights, flags, stream);
}

}  // extern "C"

template <typename T>
static int insert_bwd_impl(const void *grad_volume, const T *grad_weights, const void *slices,
                           int vol_d, int vol_h, int vol_w, const float *rot_zyx, int64_t n_rot,
                           double fftfreq_max, void *grad_slices, double *grad_rot, void *stream) {
  typedef typename Cx<T>::type C2;
  int vol_wr = vol_w / 2 + 1;
  int st = check_common(vol_d, vol_h, vol_wr, vol_d, vol_h, vol_w, n_rot, rot_zyx);
  if (st) return st;
  if (n_rot == 0 || (!grad_slices && !grad_rot)) return TT_OK;
  if (!grad_volume) return fail_invalid("grad_volume pointer is NULL");
  if (grad_rot && !slices) return fail_invalid("slices are required for the rotation gradient");
  SliceGeom g = make_geom(vol_d, vol_h, vol_w, vol_d, vol_h, vol_wr, fftfreq_max, 1e-8);
  hipStream_t s = (hipStream_t)stream;
  int64_t bpr = (g.P + kBlock - 1) / kBlock;
  int64_t rots_per_launch = std::max<int64_t>(1, kMaxBlocksPerLaunch / bpr);
  for (int64_t r0 = 0; r0 < n_rot; r0 += rots_per_launch) {
    int64_t nr = std::min(rots_per_launch, n_rot - r0);
    Launch l = plan(g, nr);
    const C2 *sl = slices ? (const C2 *)slices + r0 * g.P : nullptr;
    C2 *gs = grad_slices ? (C2 *)grad_slices + r0 * g.P : nullptr;
    double *gr = grad_rot ? grad_rot + r0 * 9 : nullptr;
    if (grad_slices && grad_rot)
      hipLaunchKernelGGL((insert_bwd_kernel<T, true, true>), dim3((unsigned)l.nb_padded),
                         dim3(kBlock), 0, s, (const C2 *)grad_volume, grad_weights, sl, g,
                         rot_zyx + r0 * 9, l.blocks_per_rot, l.n_blocks, l.nb_padded, gs, gr);
    else if (grad_slices)
      hipLaunchKernelGGL((insert_bwd_kernel<T, true, false>), dim3((unsigned)l.nb_padded),
                         dim3(kBlock), 0, s, (const C2 *)grad_volume, grad_weights, sl, g,
                         rot_zyx + r0 * 9, l.blocks_per_rot, l.n_blocks, l.nb_padded, gs, gr);
    else
      hipLaunchKernelGGL((insert_bwd_kernel<T, false, true>), dim3((unsigned)l.nb_padded),
                         dim3(kBlock), 0, s, (const C2 *)grad_volume, grad_weights, sl, g,
                         rot_zyx + r0 * 9, l.blocks_per_rot, l.n_blocks, l.nb_padded, gs, gr);
    if ((st = check_launch("insert_bwd_kernel"))) return st;
  }
  return TT_OK;
}

extern "C" {

int tt_insert_bwd_c64(const void *grad_volume, const float *grad_weights, const void *slices,
                      int vol_d, int vol_h, int vol_w, const float *rot_zyx, int64_t n_rot,
                      double fftfreq_max, void *grad_slices, double *grad_rot, void *stream) {
  return insert_bwd_impl<float>(grad_volume, grad_weights, slices, vol_d, vol_h, vol_w, rot_zyx,
                                n_rot, fftfreq_max, grad_slices, grad_rot, stream);
}

int tt_insert_bwd_c128(const void *grad_volume, const double *grad_weights, const void *slices,
                       int vol_d, int vol_h, int vol_w, const float *rot_zyx, int64_t n_rot,
                       double fftfreq_max, void *grad_slices, double *grad_rot, void *stream) {
  return insert_bwd_impl<double>(grad_volume, grad_weights, slices, vol_d, vol_h, vol_w, rot_zyx,
                                 n_rot, fftfreq_max, grad_slices, grad_rot, stream);
}

}  // extern "C"
