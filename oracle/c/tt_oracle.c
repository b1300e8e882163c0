/*
 * FP64 C restatement of the reference central-slice gather/scatter (oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * CPU-baseline leg of bench.py, never by the teamtomo_amd package.  It
 * computes exactly what oracle/fourier_slice.py computes (same coordinate
 * formulas, float32 fftfreq disc, 1e-9 integer snap), in plain C with
 * OpenMP over rotations, so that production-size problems can be checked.
 *
 * Reference (packages/primitives/torch-fourier-slice/src/torch_fourier_slice):
 *   extraction  slice_extraction/_extract_central_slices_rfft_3d.py:84-222
 *               (+ _dft_utils.py:5-72, torch-image-interpolation
 *                image_interpolation_3d.py:11-101 trilinear sampling)
 *   adjoint     the transpose of the extraction w.r.t. the volume
 *   insertion   slice_insertion/_insert_central_slices_rfft_3d.py:9-106,213-406
 * Complex arrays are interleaved (re, im) doubles.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
  int H, Wr, half_H;
  double inv_H, inv_W;
  int img_d, img_h, img_w;
  int vD, vH, vW;
  int use_mask;
  float fmax, inv_H32, inv_W32;
  double xtol, xtol10;
} geom_t;

static geom_t make_geom(int img_d, int img_h, int img_w, int vD, int vH, int vW, double fmax,
                        double xtol) {
  geom_t g;
  g.H = img_h;
  g.Wr = (int)(img_w / 2.0 + 1.0);
  g.half_H = img_h / 2;
  g.inv_H = 1.0 / img_h;
  g.inv_W = 1.0 / img_w;
  g.img_d = img_d;
  g.img_h = img_h;
  g.img_w = img_w;
  g.vD = vD;
  g.vH = vH;
  g.vW = vW;
  g.use_mask = fmax >= 0.0;
  g.fmax = (float)fmax;
  g.inv_H32 = (float)(1.0 / img_h);
  g.inv_W32 = (float)(1.0 / img_w);
  g.xtol = (double)(float)xtol;
  g.xtol10 = (double)(float)(xtol * 10.0);
  return g;
}

/* fftfreq_max disc in float32 (extraction :170-180). */
static int in_disc(const geom_t *g, int i, int j) {
  if (!g->use_mask) return 1;
  volatile float fy = (float)(i - g->half_H) * g->inv_H32;
  volatile float fx = (float)j * g->inv_W32;
  volatile float a = fy * fy, b = fx * fx;
  volatile float r2 = a + b;
  return sqrtf(r2) <= g->fmax;
}

static double snap(double c) {
  double r = rint(c);
  return fabs(c - r) < 1e-9 ? r : c;
}

/* Rotated, folded array coordinate of pixel (i, j); returns conj flag.
 * insert != 0: conj iff x' < 0; else the extraction tolerance rule. */
static int coord(const geom_t *g, const float *R, int i, int j, int insert, double c[3]) {
  double fy = (double)(i - g->half_H) / g->H;
  double fx = (double)j / g->img_w;
  double q[3];
  for (int d = 0; d < 3; ++d) q[d] = (double)R[3 * d + 1] * fy + (double)R[3 * d + 2] * fx;
  int conj;
  if (insert) {
    conj = q[2] < 0.0;
  } else {
    double r11 = R[8], r12 = R[7];
    int tie = fabs(r11) <= g->xtol10 ? (r12 < 0.0) : (r11 < 0.0);
    conj = (q[2] < -g->xtol) || (fabs(q[2]) < g->xtol && tie);
  }
  if (conj)
    for (int d = 0; d < 3; ++d) q[d] = -q[d];
  c[0] = snap((q[0] + 0.5) / (1.0 - 1.0 / g->img_d) * (g->img_d - 1));
  c[1] = snap((q[1] + 0.5) / (1.0 - 1.0 / g->img_h) * (g->img_h - 1));
  c[2] = snap((q[2] * 2.0) * (g->Wr - 1));
  return conj;
}

static int inside(const geom_t *g, const double c[3]) {
  return c[0] >= 0 && c[0] <= g->vD - 1 && c[1] >= 0 && c[1] <= g->vH - 1 && c[2] >= 0 &&
         c[2] <= g->vW - 1;
}

int tt_oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void tt_oracle_extract(const double *vol, int vD, int vH, int vW, int img_d, int img_h, int img_w,
                       const float *rot, int64_t n, double fmax, double xtol, double *out) {
  geom_t g = make_geom(img_d, img_h, img_w, vD, vH, vW, fmax, xtol);
  const int64_t P = (int64_t)g.H * g.Wr;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t r = 0; r < n; ++r) {
    const float *R = rot + 9 * r;
    for (int i = 0; i < g.H; ++i) {
      for (int j = 0; j < g.Wr; ++j) {
        double *o = out + 2 * (r * P + (int64_t)i * g.Wr + j);
        double c[3];
        int conj = coord(&g, R, i, j, 0, c);
        o[0] = o[1] = 0.0;
        if (!in_disc(&g, i, j) || !inside(&g, c)) continue;
        int i0[3];
        double t[3];
        for (int d = 0; d < 3; ++d) {
          i0[d] = (int)floor(c[d]);
          t[d] = c[d] - i0[d];
        }
        const int N[3] = {vD, vH, vW};
        double re = 0, im = 0;
        for (int q = 0; q < 8; ++q) {
          int k[3] = {i0[0] + (q >> 2), i0[1] + ((q >> 1) & 1), i0[2] + (q & 1)};
          if (k[0] >= N[0] || k[1] >= N[1] || k[2] >= N[2]) continue;
          double w = ((q >> 2) ? t[0] : 1 - t[0]) * (((q >> 1) & 1) ? t[1] : 1 - t[1]) *
                     ((q & 1) ? t[2] : 1 - t[2]);
          const double *v = vol + 2 * (((int64_t)k[0] * vH + k[1]) * vW + k[2]);
          re += w * v[0];
          im += w * v[1];
        }
        o[0] = re;
        o[1] = conj ? -im : im;
      }
    }
  }
}

/* gV += A^T gS (extraction adjoint).  Per-thread partial volumes are summed
 * in thread order, so the result does not depend on scheduling. */
void tt_oracle_extract_adjoint(const double *gS, int vD, int vH, int vW, int img_d, int img_h,
                               int img_w, const float *rot, int64_t n, double fmax, double xtol,
                               double *gV) {
  geom_t g = make_geom(img_d, img_h, img_w, vD, vH, vW, fmax, xtol);
  const int64_t P = (int64_t)g.H * g.Wr, NV = (int64_t)vD * vH * vW;
  int nt = tt_oracle_threads();
  double *part = (double *)calloc((size_t)nt * 2 * NV, sizeof(double));
  if (!part) return;
#pragma omp parallel
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    double *acc = part + (size_t)tid * 2 * NV;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
      const float *R = rot + 9 * r;
      for (int i = 0; i < g.H; ++i)
        for (int j = 0; j < g.Wr; ++j) {
          double c[3];
          int conj = coord(&g, R, i, j, 0, c);
          if (!in_disc(&g, i, j) || !inside(&g, c)) continue;
          const double *s = gS + 2 * (r * P + (int64_t)i * g.Wr + j);
          double gr = s[0], gi = conj ? -s[1] : s[1];
          int i0[3];
          double t[3];
          for (int d = 0; d < 3; ++d) {
            i0[d] = (int)floor(c[d]);
            t[d] = c[d] - i0[d];
          }
          const int N[3] = {vD, vH, vW};
          for (int q = 0; q < 8; ++q) {
            int k[3] = {i0[0] + (q >> 2), i0[1] + ((q >> 1) & 1), i0[2] + (q & 1)};
            if (k[0] >= N[0] || k[1] >= N[1] || k[2] >= N[2]) continue;
            double w = ((q >> 2) ? t[0] : 1 - t[0]) * (((q >> 1) & 1) ? t[1] : 1 - t[1]) *
                       ((q & 1) ? t[2] : 1 - t[2]);
            double *v = acc + 2 * (((int64_t)k[0] * vH + k[1]) * vW + k[2]);
            v[0] += w * gr;
            v[1] += w * gi;
          }
        }
    }
  }
  for (int t = 0; t < nt; ++t)
    for (int64_t e = 0; e < 2 * NV; ++e) gV[e] += part[(size_t)t * 2 * NV + e];
  free(part);
}

/* (vol, w) += insertion, floor/ceil corners and x=0 Friedel mirror. */
void tt_oracle_insert(const double *slices, int D, int H, int W, const float *rot, int64_t n,
                      double fmax, double *vol, double *wts) {
  const int Wr = (int)(W / 2.0 + 1.0);
  geom_t g = make_geom(D, H, W, D, H, Wr, fmax, 1e-8);
  const int64_t P = (int64_t)H * Wr, NV = (int64_t)D * H * Wr;
  int nt = tt_oracle_threads();
  double *part = (double *)calloc((size_t)nt * 3 * NV, sizeof(double));
  if (!part) return;
#pragma omp parallel
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
#else
    int tid = 0;
#endif
    double *av = part + (size_t)tid * 3 * NV, *aw = av + 2 * NV;
#pragma omp for schedule(static)
    for (int64_t r = 0; r < n; ++r) {
      const float *R = rot + 9 * r;
      for (int i = 0; i < H; ++i)
        for (int j = 0; j < Wr; ++j) {
          if (j == 0 && i < H / 2) continue;
          double c[3];
          int conj = coord(&g, R, i, j, 1, c);
          if (!in_disc(&g, i, j) || !inside(&g, c)) continue;
          const double *s = slices + 2 * (r * P + (int64_t)i * Wr + j);
          double dr = s[0], di = conj ? -s[1] : s[1];
          int lo[3], hi[3];
          double t[3];
          for (int d = 0; d < 3; ++d) {
            lo[d] = (int)floor(c[d]);
            hi[d] = (int)ceil(c[d]);
            t[d] = c[d] - lo[d];
          }
          for (int q = 0; q < 8; ++q) {
            int dz = q >> 2, dy = (q >> 1) & 1, dx = q & 1;
            int z = dz ? hi[0] : lo[0], y = dy ? hi[1] : lo[1], x = dx ? hi[2] : lo[2];
            double w = (dz ? t[0] : 1 - t[0]) * (dy ? t[1] : 1 - t[1]) * (dx ? t[2] : 1 - t[2]);
            int64_t e = ((int64_t)z * H + y) * Wr + x;
            av[2 * e] += w * dr;
            av[2 * e + 1] += w * di;
            aw[e] += w;
            if (x == 0 && !(z == D / 2 && y == H / 2)) {
              int64_t m = ((int64_t)((D - z) % D) * H + (H - y) % H) * Wr;
              av[2 * m] += w * dr;
              av[2 * m + 1] -= w * di;
              aw[m] += w;
            }
          }
        }
    }
  }
  for (int t = 0; t < nt; ++t) {
    const double *av = part + (size_t)t * 3 * NV, *aw = av + 2 * NV;
    for (int64_t e = 0; e < 2 * NV; ++e) vol[e] += av[e];
    for (int64_t e = 0; e < NV; ++e) wts[e] += aw[e];
  }
  free(part);
}
