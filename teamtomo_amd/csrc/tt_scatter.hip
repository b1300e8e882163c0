// Pixel-driven scatter kernels: one lane per slice pixel, 8 trilinear corners
// accumulated with hardware float atomics in L2 (global_atomic_add_f32 /
// _f64, no CAS loop).  Used for the extraction adjoint and the insertion.
//
// Insertion semantics follow packages/primitives/torch-fourier-slice/src/
// torch_fourier_slice/slice_insertion/_insert_central_slices_rfft_3d.py:
//   :40-48  skip slice pixels with fx == 0 and fy < 0
//   :78-80  conjugate where x' < 0 (strict)
//   :271-275 drop points outside [0, N-1]
//   :298-346 floor/ceil corners, volume += w d, weights += w
//   :348-404 x=0-plane Friedel mirror at ((D-z)%D, (H-y)%H, 0), DC excluded
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_engine.h"
#include "tt_scatter.h"

namespace tt {

namespace {

constexpr int kBlock = 256;
const int64_t kMaxBlocksPerLaunch = (int64_t)1 << 23;

__device__ __forceinline__ int64_t logical_block_s(int64_t nb_padded) {
  int64_t b = blockIdx.x;
  int64_t per = nb_padded >> 3;
  return (b & 7) * per + (b >> 3);
}

__global__ __launch_bounds__(kBlock) void adjoint_atomic_kernel(
    const f2v *__restrict__ gS, SliceGeom g, const float *__restrict__ rot, int blocks_per_rot,
    int64_t n_blocks, int64_t nb_padded, f2v *__restrict__ gV) {
  int64_t lb = logical_block_s(nb_padded);
  if (lb >= n_blocks) return;
  int64_t r_idx = lb / blocks_per_rot;
  int p = (int)(lb - r_idx * blocks_per_rot) * kBlock + (int)threadIdx.x;
  if (p >= g.P) return;
  Rot r = load_rot(rot + r_idx * 9, g);
  int i, j;
  split_pixel(g, p, i, j);
  Coord c = slice_coord<false>(g, r, i, j);
  if (!(in_disc(g, i, j) && inside(g, c))) return;
  Cell k = cell_of(g, c);
  f2v gs = gS[r_idx * g.P + p];
  if (c.conj) gs.y = -gs.y;
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  const int64_t b0 = k.z0 * sz + k.y0 * sy + k.x0;
  float ux = 1.0f - k.tx, uy = 1.0f - k.ty, uz = 1.0f - k.tz;
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

template <typename T>
__global__ __launch_bounds__(kBlock) void insert_atomic_kernel(
    const typename Cx<T>::type *__restrict__ slices, SliceGeom g, const float *__restrict__ rot,
    int blocks_per_rot, int64_t n_blocks, int64_t nb_padded, T *__restrict__ vol,
    T *__restrict__ wts) {
  int64_t lb = logical_block_s(nb_padded);
  if (lb >= n_blocks) return;
  int64_t r_idx = lb / blocks_per_rot;
  int p = (int)(lb - r_idx * blocks_per_rot) * kBlock + (int)threadIdx.x;
  if (p >= g.P) return;
  Rot r = load_rot(rot + r_idx * 9, g);
  int i, j;
  split_pixel(g, p, i, j);
  if (j == 0 && i < g.half_H) return;
  Coord c = slice_coord<true>(g, r, i, j);
  if (!(in_disc(g, i, j) && inside(g, c))) return;
  Cell k = cell_of(g, c);
  typename Cx<T>::type d = slices[r_idx * g.P + p];
  if (c.conj) d.y = -d.y;
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  T tz = (T)k.tz, ty = (T)k.ty, tx = (T)k.tx;
  T uz = (T)1 - tz, uy = (T)1 - ty, ux = (T)1 - tx;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
    int z = k.z0 + dz, y = k.y0 + dy, x = k.x0 + dx;
    T w = (dz ? tz : uz) * (dy ? ty : uy) * (dx ? tx : ux);
    int64_t e = (int64_t)z * sz + (int64_t)y * sy + x;
    unsafeAtomicAdd(vol + 2 * e, w * d.x);
    unsafeAtomicAdd(vol + 2 * e + 1, w * d.y);
    unsafeAtomicAdd(wts + e, w);
    if (x == 0 && !(z == g.vD / 2 && y == g.vH / 2)) {
      int64_t m = (int64_t)((g.vD - z) % g.vD) * sz + (int64_t)((g.vH - y) % g.vH) * sy;
      unsafeAtomicAdd(vol + 2 * m, w * d.x);
      unsafeAtomicAdd(vol + 2 * m + 1, -w * d.y);
      unsafeAtomicAdd(wts + m, w);
    }
  }
}

struct Plan {
  int blocks_per_rot;
  int64_t rots_per_launch;
};

Plan make_plan(const SliceGeom &g) {
  Plan p;
  p.blocks_per_rot = (g.P + kBlock - 1) / kBlock;
  p.rots_per_launch = std::max<int64_t>(1, kMaxBlocksPerLaunch / p.blocks_per_rot);
  return p;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void zero_kernel(T *__restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    p[i] = (T)0;
}

}  // namespace

template <typename T> int zero_fill(T *p, int64_t n, hipStream_t s) {
  if (n <= 0) return TT_OK;
  int grid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 8192);
  hipLaunchKernelGGL(zero_kernel<T>, dim3(grid), dim3(kBlock), 0, s, p, n);
  return check_launch("zero_kernel");
}
template int zero_fill<float>(float *, int64_t, hipStream_t);
template int zero_fill<double>(double *, int64_t, hipStream_t);

int scatter_extract_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot,
                            f2v *gV, int flags, hipStream_t s) {
  const bool det = flags & TT_FLAG_DETERMINISTIC, overwrite = flags & TT_FLAG_OVERWRITE;
  // The brick engine accumulates in fixed point: deterministic as it stands.
  if (!(flags & TT_FLAG_PIXEL_ATOMIC) && engine_supports(g, n_rot))
    return engine_adjoint(g, gS, rot, n_rot, gV, !overwrite, s);
  if (det) return fail_invalid("the pixel-atomic scatter is not deterministic (and the geometry is beyond "
                               "the brick engine's limits, or TT_FLAG_PIXEL_ATOMIC was given)");
  int st;
  if (overwrite && (st = zero_fill<float>((float *)gV, 2 * (int64_t)g.vD * g.vH * g.vW, s))) return st;
  Plan pl = make_plan(g);
  for (int64_t r0 = 0; r0 < n_rot; r0 += pl.rots_per_launch) {
    int64_t nr = std::min(pl.rots_per_launch, n_rot - r0);
    int64_t nb = (int64_t)pl.blocks_per_rot * nr, nbp = (nb + 7) & ~(int64_t)7;
    hipLaunchKernelGGL(adjoint_atomic_kernel, dim3((unsigned)nbp), dim3(kBlock), 0, s,
                       gS + r0 * g.P, g, rot + r0 * 9, pl.blocks_per_rot, nb, nbp, gV);
    if ((st = check_launch("adjoint_atomic_kernel"))) return st;
  }
  return TT_OK;
}

template <typename T>
int scatter_insert(const SliceGeom &g, const typename Cx<T>::type *slices, const float *rot,
                   int64_t n_rot, typename Cx<T>::type *vol, T *w, int flags, hipStream_t s) {
  const bool det = flags & TT_FLAG_DETERMINISTIC, overwrite = flags & TT_FLAG_OVERWRITE;
  if (sizeof(T) == sizeof(float) && !(flags & TT_FLAG_PIXEL_ATOMIC) && engine_supports(g, n_rot))
    return engine_insert(g, (const f2v *)slices, rot, n_rot, (f2v *)vol, (float *)w, !overwrite, s);
  if (det) return fail_invalid("deterministic insertion is available for complex64 within the brick "
                               "engine's limits only");
  int st;
  const int64_t nv = (int64_t)g.vD * g.vH * g.vW;
  if (overwrite && ((st = zero_fill<T>((T *)vol, 2 * nv, s)) || (st = zero_fill<T>(w, nv, s))))
    return st;
  Plan pl = make_plan(g);
  for (int64_t r0 = 0; r0 < n_rot; r0 += pl.rots_per_launch) {
    int64_t nr = std::min(pl.rots_per_launch, n_rot - r0);
    int64_t nb = (int64_t)pl.blocks_per_rot * nr, nbp = (nb + 7) & ~(int64_t)7;
    hipLaunchKernelGGL((insert_atomic_kernel<T>), dim3((unsigned)nbp), dim3(kBlock), 0, s,
                       slices + r0 * g.P, g, rot + r0 * 9, pl.blocks_per_rot, nb, nbp,
                       reinterpret_cast<T *>(vol), w);
    if ((st = check_launch("insert_atomic_kernel"))) return st;
  }
  return TT_OK;
}

template int scatter_insert<float>(const SliceGeom &, const f2v *, const float *, int64_t, f2v *,
                                   float *, int, hipStream_t);
template int scatter_insert<double>(const SliceGeom &, const d2v *, const float *, int64_t, d2v *,
                                    double *, int, hipStream_t);

}  // namespace tt
