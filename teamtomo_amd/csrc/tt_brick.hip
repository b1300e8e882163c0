// Brick-stationary scatter for gfx950: the extraction adjoint and the
// central-slice insertion without global atomics.
//
// Each workgroup owns one BZ x BY x BX brick of the output volume and keeps
// its accumulator in LDS.  It (1) tests every rotation's slice plane against
// the brick (one lane per rotation), (2) for each crossing (rotation, Friedel
// half) enumerates exactly the slice pixels whose trilinear footprint
// overlaps the brick -- one lane per slice row computes the row's pixel
// interval, a wave prefix sum flattens the rows and every lane takes one
// pixel -- and adds w * value into LDS with ds_add_f32, and (3) writes each
// voxel once.  A pixel whose footprint straddles a brick face is visited by
// both bricks, each adding only its own corners, so every voxel sees exactly
// the contributions of the pixel-driven scatter.
//
// Insertion semantics (packages/primitives/torch-fourier-slice/src/
// torch_fourier_slice/slice_insertion/_insert_central_slices_rfft_3d.py):
// the direct splat goes to the bricks; the x=0-plane Friedel mirror (:348-404)
// is applied afterwards from a scratch copy of the plane:
//   V[z,y,0] = P[z,y] + conj(P[mz,my])   unless (mz,my) is DC,
// with (mz,my) = ((D-z)%D, (H-y)%H), because the mirror of every x=0 corner
// other than DC is added conjugated at that point.
//
// Deterministic mode accumulates in 64-bit fixed point (integer adds commute,
// so the result is bitwise reproducible); the scale 2^k is chosen on the
// device from max|value| so that no partial sum can overflow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/teamtomo_amd.h"
#include "tt_common.h"
#include "tt_scatter.h"

namespace tt {

namespace {

constexpr int kT = 256;
constexpr int kWaves = kT / 64;
constexpr int kRotChunk = kT;
constexpr int kItemsMax = 2 * kRotChunk;
constexpr double kMargin = 1e-6;

__device__ __forceinline__ int64_t logical_block_b(int64_t nb_padded) {
  int64_t b = blockIdx.x;
  int64_t per = nb_padded >> 3;
  return (b & 7) * per + (b >> 3);
}

// Accumulator element for float (fast) and fixed-point (deterministic) modes.
template <bool DET> struct Acc;
template <> struct Acc<false> {
  typedef float T;
  __device__ static void add(T *p, float v, double) { atomicAdd(p, v); }
  __device__ static float get(T v, double) { return v; }
};
template <> struct Acc<true> {
  typedef unsigned long long T;
  __device__ static void add(T *p, float v, double scale) {
    long long q = (long long)rint((double)v * scale);
    atomicAdd(p, (unsigned long long)q);
  }
  __device__ static float get(T v, double scale) { return (float)((double)(long long)v / scale); }
};

// Per-rotation geometry in array-coordinate space: c = beta + s (A m + B j)
// with m = i - H/2.
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

// Scale exponent for fixed-point accumulation: every partial sum is bounded
// by max|v| * n_rot * per_slice (per_slice bounds the summed weights one
// slice can put on one voxel), kept below 2^62.
__device__ __forceinline__ double fixed_scale(float maxabs, double bound) {
  double m = (double)maxabs * bound;
  if (!(m > 0.0)) return 1.0;
  int e;
  frexp(m, &e);  // m < 2^e
  return ldexp(1.0, 62 - e);
}

template <int BZ, int BY, int BX, bool INSERT, bool DET>
__global__ __launch_bounds__(kT) void brick_scatter_kernel(
    const f2v *__restrict__ src, SliceGeom g, const float *__restrict__ rot, int n_rot, int nby,
    int nbx, int n_bricks, int nb_padded, const float *__restrict__ maxabs, double bound_v,
    double bound_w, f2v *__restrict__ out, float *__restrict__ wout, f2v *__restrict__ plane,
    float *__restrict__ wplane, int accumulate) {
  typedef Acc<DET> A;
  constexpr int NV = BZ * BY * BX;
  __shared__ typename A::T s_re[NV];
  __shared__ typename A::T s_im[NV];
  __shared__ typename A::T s_w[INSERT ? NV : 1];
  __shared__ int s_item[kItemsMax];
  __shared__ int s_m0[kItemsMax];
  __shared__ int s_m1[kItemsMax];
  __shared__ int s_nitems;

  const int64_t lb = logical_block_b(nb_padded);
  if (lb >= n_bricks) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bz = (int)(lb / ((int64_t)nby * nbx));
  const int rem = (int)(lb - (int64_t)bz * nby * nbx);
  const int by = rem / nbx, bx = rem - by * nbx;
  const int lo[3] = {bz * BZ, by * BY, bx * BX};
  const int hi[3] = {min(lo[0] + BZ - 1, g.vD - 1), min(lo[1] + BY - 1, g.vH - 1),
                     min(lo[2] + BX - 1, g.vW - 1)};
  // extended box of coordinates whose 2x2x2 footprint can touch the brick
  double L[3], U[3], ctr[3], half[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    L[d] = (double)lo[d] - 1.0 - kMargin;
    U[d] = (double)hi[d] + 1.0 + kMargin;
    ctr[d] = 0.5 * (L[d] + U[d]);
    half[d] = 0.5 * (U[d] - L[d]);
  }
  const double beta[3] = {g.bz, g.by, 0.0};
  double scale_v = 1.0, scale_w = 1.0;
  if (DET) {
    scale_v = fixed_scale(*maxabs, bound_v);
    scale_w = fixed_scale(1.0f, bound_w);
  }

  for (int v = tid; v < NV; v += kT) {
    s_re[v] = 0;
    s_im[v] = 0;
    if (INSERT) s_w[v] = 0;
  }

  for (int r0 = 0; r0 < n_rot; r0 += kRotChunk) {
    if (tid == 0) s_nitems = 0;
    __syncthreads();
    // (1) plane / brick intersection, one lane per rotation
    const int r = r0 + tid;
    if (r < n_rot) {
      PlaneGeom p = plane_geom(g, rot + (int64_t)r * 9);
      double n0 = p.A[1] * p.B[2] - p.A[2] * p.B[1];
      double n1 = p.A[2] * p.B[0] - p.A[0] * p.B[2];
      double n2 = p.A[0] * p.B[1] - p.A[1] * p.B[0];
      double dist = n0 * (ctr[0] - beta[0]) + n1 * (ctr[1] - beta[1]) + n2 * (ctr[2] - beta[2]);
      double rad = fabs(n0) * half[0] + fabs(n1) * half[1] + fabs(n2) * half[2];
      double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
      if (fabs(dist) <= rad + 1e-9 * nn) {
        double AA = p.A[0] * p.A[0] + p.A[1] * p.A[1] + p.A[2] * p.A[2];
        double BB = p.B[0] * p.B[0] + p.B[1] * p.B[1] + p.B[2] * p.B[2];
        double AB = p.A[0] * p.B[0] + p.A[1] * p.B[1] + p.A[2] * p.B[2];
        double det = AA * BB - AB * AB;
        double mc = 0.0, mr = 0.0, jc = 0.0, jr = 0.0;
        if (det > 1e-20 * AA * BB) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double pm = (BB * p.A[d] - AB * p.B[d]) / det;
            double pj = (AA * p.B[d] - AB * p.A[d]) / det;
            mc += pm * (ctr[d] - beta[d]);
            jc += pj * (ctr[d] - beta[d]);
            mr += fabs(pm) * half[d];
            jr += fabs(pj) * half[d];
          }
        } else {  // degenerate (non-rotation) matrix: search every row and column
          mr = jr = 1e300;
        }
        const double mmin = (double)(-g.half_H), mmax = (double)(g.H - 1 - g.half_H);
        for (int sh = 0; sh < 2; ++sh) {
          double s = sh ? -1.0 : 1.0;
          double jlo = s * jc - jr, jhi = s * jc + jr;
          if (jhi < -kMargin || jlo > (double)(g.Wr - 1) + kMargin) continue;
          double mlo = fmax(s * mc - mr, mmin), mhi = fmin(s * mc + mr, mmax);
          int m0 = (int)ceil(mlo - kMargin), m1 = (int)floor(mhi + kMargin);
          if (m0 > m1) continue;
          int slot = atomicAdd(&s_nitems, 1);
          s_item[slot] = 2 * r + sh;
          s_m0[slot] = m0;
          s_m1[slot] = m1;
        }
      }
    }
    __syncthreads();
    const int nitems = s_nitems;
    // (2) each wave walks its items row batch by row batch
    for (int it = wave; it < nitems; it += kWaves) {
      const int code = s_item[it];
      const int ri = code >> 1;
      const double s = (code & 1) ? -1.0 : 1.0;
      const int m0 = s_m0[it], m1 = s_m1[it];
      const float *Rp = rot + (int64_t)ri * 9;
      PlaneGeom p = plane_geom(g, Rp);
      Rot rr = load_rot(Rp, g);
      double C1[3], inv1[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        C1[d] = s * p.B[d];
        inv1[d] = fabs(C1[d]) > 1e-12 ? 1.0 / C1[d] : 0.0;
      }
      for (int mb = m0; mb <= m1; mb += 64) {
        const int m = mb + lane;
        int jl = 0, len = 0;
        if (m <= m1) {
          double jmin = 0.0, jmax = (double)(g.Wr - 1);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            double c0 = beta[d] + s * p.A[d] * (double)m;
            if (inv1[d] != 0.0) {
              double t1 = (L[d] - c0) * inv1[d], t2 = (U[d] - c0) * inv1[d];
              jmin = fmax(jmin, fmin(t1, t2));
              jmax = fmin(jmax, fmax(t1, t2));
            } else if (c0 < L[d] || c0 > U[d]) {
              jmax = -1.0;
            }
          }
          jl = (int)ceil(jmin - kMargin);
          int jh = (int)floor(jmax + kMargin);
          jl = max(jl, 0);
          jh = min(jh, g.Wr - 1);
          len = max(0, jh - jl + 1);
        }
        // inclusive wave prefix sum of row lengths
        int incl = len;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          int t = __shfl_up(incl, off, 64);
          if (lane >= off) incl += t;
        }
        const int total = __shfl(incl, 63, 64);
        for (int tb = 0; tb < total; tb += 64) {
          const int t = tb + lane;
          // row k = number of rows whose inclusive prefix is <= t
          int k = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1) {
            int v = __shfl(incl, k + step - 1, 64);
            if (v <= t) k += step;
          }
          const int row_incl = __shfl(incl, k, 64);
          const int row_jl = __shfl(jl, k, 64);
          const int row_len = __shfl(len, k, 64);
          if (t >= total) continue;
          const int j = row_jl + (t - (row_incl - row_len));
          const int i = mb + k + g.half_H;
          Coord c = slice_coord<INSERT>(g, rr, i, j);
          if (c.conj != (s < 0.0)) continue;
          if (INSERT && j == 0 && i < g.half_H) continue;
          if (!in_disc(g, i, j) || !inside(g, c)) continue;
          Cell k3 = cell_of(g, c);
          const int z0 = k3.z0, y0 = k3.y0, x0 = k3.x0;
          if (z0 > hi[0] || z0 + 1 < lo[0] || y0 > hi[1] || y0 + 1 < lo[1] || x0 > hi[2] ||
              x0 + 1 < lo[2])
            continue;
          f2v v = src[(int64_t)ri * g.P + (int64_t)i * g.Wr + j];
          if (c.conj) v.y = -v.y;
          const float wz[2] = {1.0f - k3.tz, k3.tz};
          const float wy[2] = {1.0f - k3.ty, k3.ty};
          const float wx[2] = {1.0f - k3.tx, k3.tx};
#pragma unroll
          for (int dz = 0; dz < 2; ++dz) {
            const int z = z0 + dz;
            if (z < lo[0] || z > hi[0]) continue;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
              const int y = y0 + dy;
              if (y < lo[1] || y > hi[1]) continue;
              const float wzy = wz[dz] * wy[dy];
#pragma unroll
              for (int dx = 0; dx < 2; ++dx) {
                const int x = x0 + dx;
                if (x < lo[2] || x > hi[2]) continue;
                const float w = wzy * wx[dx];
                const int e = ((z - lo[0]) * BY + (y - lo[1])) * BX + (x - lo[2]);
                A::add(&s_re[e], w * v.x, scale_v);
                A::add(&s_im[e], w * v.y, scale_v);
                if (INSERT) A::add(&s_w[e], w, scale_w);
              }
            }
          }
        }
      }
    }
    __syncthreads();
  }

  // (3) flush: every voxel of the brick is written by this block only
  const int64_t sy = g.vW, sz = (int64_t)g.vH * g.vW;
  for (int v = tid; v < NV; v += kT) {
    const int lz = v / (BY * BX), ly = (v / BX) % BY, lx = v % BX;
    const int z = lo[0] + lz, y = lo[1] + ly, x = lo[2] + lx;
    if (z > hi[0] || y > hi[1] || x > hi[2]) continue;
    f2v val = {A::get(s_re[v], scale_v), A::get(s_im[v], scale_v)};
    if (INSERT && x == 0) {
      plane[(int64_t)z * g.vH + y] = val;
      wplane[(int64_t)z * g.vH + y] = A::get(s_w[v], scale_w);
      continue;
    }
    const int64_t e = (int64_t)z * sz + (int64_t)y * sy + x;
    if (accumulate) val += out[e];
    out[e] = val;
    if (INSERT) wout[e] = A::get(s_w[v], scale_w) + (accumulate ? wout[e] : 0.0f);
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
    f2v m = plane[mz * H + my];
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
__global__ __launch_bounds__(kT) void maxabs_kernel(const f2v *__restrict__ x, int64_t n,
                                                    unsigned int *__restrict__ res) {
  float m = 0.0f;
  for (int64_t p = (int64_t)blockIdx.x * kT + threadIdx.x; p < n; p += (int64_t)gridDim.x * kT) {
    f2v v = x[p];
    m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
  }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(res, __float_as_uint(m));
}

__global__ void zero_u32_kernel(unsigned int *p) { *p = 0u; }

template <int BZ, int BY, int BX, bool INSERT, bool DET>
int launch_brick(const SliceGeom &g, const f2v *src, const float *rot, int64_t n_rot, f2v *out,
                 float *wout, int accumulate, hipStream_t s) {
  const int nbz = (g.vD + BZ - 1) / BZ, nby = (g.vH + BY - 1) / BY, nbx = (g.vW + BX - 1) / BX;
  const int nb = nbz * nby * nbx;
  const int nbp = (nb + 7) & ~7;
  float *maxabs = nullptr;
  f2v *plane = nullptr;
  float *wplane = nullptr;
  void *scratch = nullptr;
  const size_t plane_elems = (size_t)g.vD * g.vH;
  size_t bytes = 256;  // maxabs slot
  if (INSERT) bytes += plane_elems * (sizeof(f2v) + sizeof(float));
  if (DET || INSERT) {
    hipError_t e = hipMallocAsync(&scratch, bytes, s);
    if (e != hipSuccess) {
      set_error(std::string("hipMallocAsync: ") + hipGetErrorString(e));
      return TT_ERR_HIP;
    }
    maxabs = (float *)scratch;
    if (INSERT) {
      plane = (f2v *)((char *)scratch + 256);
      wplane = (float *)((char *)scratch + 256 + plane_elems * sizeof(f2v));
    }
  }
  // per-slice bound on the summed trilinear weights one slice puts on one voxel
  const double step_y = (double)g.ay * g.inv_H, step_x = (double)g.ax * g.inv_W;
  const double per_px = std::max(step_y, step_x);
  const double per_slice = std::pow(2.0 / std::max(per_px, 1e-3) + 2.0, 2.0);
  const double bound = (double)n_rot * per_slice + 1.0;
  if (DET) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, (unsigned int *)maxabs);
    int64_t n = n_rot * (int64_t)g.P;
    int grid = (int)std::min<int64_t>((n + kT - 1) / kT, 4096);
    hipLaunchKernelGGL(maxabs_kernel, dim3(std::max(grid, 1)), dim3(kT), 0, s, src, n,
                       (unsigned int *)maxabs);
  }
  hipLaunchKernelGGL((brick_scatter_kernel<BZ, BY, BX, INSERT, DET>), dim3(nbp), dim3(kT), 0, s,
                     src, g, rot, (int)n_rot, nby, nbx, nb, nbp, maxabs, bound, bound, out, wout,
                     plane, wplane, accumulate);
  int st = check_launch("brick_scatter_kernel");
  if (!st && INSERT) {
    int n = (int)plane_elems;
    hipLaunchKernelGGL(mirror_plane_kernel, dim3((n + kT - 1) / kT), dim3(kT), 0, s, plane, wplane,
                       g.vD, g.vH, g.vW, out, wout, accumulate);
    st = check_launch("mirror_plane_kernel");
  }
  if (scratch) {
    hipError_t e = hipFreeAsync(scratch, s);
    if (!st && e != hipSuccess) {
      set_error(std::string("hipFreeAsync: ") + hipGetErrorString(e));
      st = TT_ERR_HIP;
    }
  }
  return st;
}

}  // namespace

int brick_scatter_adjoint(const SliceGeom &g, const f2v *gS, const float *rot, int64_t n_rot,
                          f2v *gV, int deterministic, int accumulate, hipStream_t s) {
  if (deterministic)
    return launch_brick<16, 16, 8, false, true>(g, gS, rot, n_rot, gV, nullptr, accumulate, s);
  return launch_brick<16, 16, 16, false, false>(g, gS, rot, n_rot, gV, nullptr, accumulate, s);
}

int brick_scatter_insert(const SliceGeom &g, const f2v *slices, const float *rot, int64_t n_rot,
                         f2v *vol, float *w, int deterministic, int accumulate, hipStream_t s) {
  if (deterministic)
    return launch_brick<16, 16, 8, true, true>(g, slices, rot, n_rot, vol, w, accumulate, s);
  return launch_brick<16, 16, 16, true, false>(g, slices, rot, n_rot, vol, w, accumulate, s);
}

}  // namespace tt
