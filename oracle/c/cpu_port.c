/* CPU port of the batched propagation path -- TEST AND CPU-BASELINE INFRASTRUCTURE.
 *
 * OpenMP C restatements timed as bench.py's cpu_baseline leg:
 *   plain_run      Lens(z=0.5, f) + Detector(z=1.0) run_to_end + binning, f32
 *                  operation order of reference propagator.py:51-72 and
 *                  components.py:161-174, binning of grid.py:120-153 and
 *                  utils.py:83-114 (per-thread private histograms).
 *   stem_fit       loss and gradient of the 4D-STEM descan fit in f64:
 *                  per (scan position, camera length) residual moments, then
 *                  the same parameter chain rule as the GPU kernel.
 * Built with -O2 -fopenmp -ffp-contract=off (no FMA contraction).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

int cpu_threads(void) { return omp_get_max_threads(); }

static int32_t xla_i32(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return -2147483647 - 1;
  return (int32_t)v;
}

void plain_run(int64_t n, const float* x, const float* y, const float* dx, const float* dy, double f,
               const double* m /* 6 */, int64_t h, int64_t w, float* ox, float* oy, float* odx, float* ody,
               float* opl, int64_t* hist) {
  const float d1 = (float)0.5, d2 = (float)0.5, fl = (float)f, f2 = (float)(2.0 * f);
  const float pl0 = (float)0.5;
#pragma omp parallel
  {
    int64_t* priv = (int64_t*)calloc((size_t)(h * w), sizeof(int64_t));
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      float xi = x[i] + dx[i] * d1, yi = y[i] + dy[i] * d1;
      float ndx = -xi / fl + dx[i], ndy = -yi / fl + dy[i];
      float pl = pl0 - (xi * xi + yi * yi) / f2;
      xi = xi + ndx * d2;
      yi = yi + ndy * d2;
      pl = pl + d2;
      ox[i] = xi; oy[i] = yi; odx[i] = ndx; ody[i] = ndy; opl[i] = pl;
      double fy = (m[0] * (double)yi + m[1] * (double)xi) + m[2];
      double fx = (m[3] * (double)yi + m[4] * (double)xi) + m[5];
      int32_t iy = xla_i32(nearbyint(fy)), ix = xla_i32(nearbyint(fx));
      if (iy >= 0 && iy < h && ix >= 0 && ix < w) priv[iy * w + ix] += 1;
    }
#pragma omp critical
    for (int64_t k = 0; k < h * w; ++k) hist[k] += priv[k];
    free(priv);
  }
}

void stem_fit(int scan_h, int scan_w, int n_side, int n_cam, double semi_conv, double defocus, double pivot,
              const double* off, const double* P, const float* T, int64_t s_begin, int64_t s_end, double* out) {
  const int64_t S = (int64_t)scan_h * scan_w, J = (int64_t)n_side * n_side;
  const int np = 18 + n_cam + 1;
  double* th = (double*)malloc(sizeof(double) * n_side);
  for (int k = 0; k < n_side; ++k)
    th[k] = (double)(float)(((double)k - 0.5 * (n_side - 1)) * (2.0 * semi_conv / n_side));
  const double* A = P;
  const double pi = 3.14159265358979323846;
  const double phi = P[17] * (pi / 180.0);
  memset(out, 0, sizeof(double) * np);
#pragma omp parallel
  {
    double* g = (double*)calloc(np, sizeof(double));
#pragma omp for schedule(static)
    for (int64_t s = s_begin; s < s_end; ++s) {
      const int64_t i = s / scan_w, j = s % scan_w;
      const double u = P[16] * ((double)i - 0.5 * (scan_h - 1)), v = P[16] * ((double)j - 0.5 * (scan_w - 1));
      const double spx = -sin(phi) * u + cos(phi) * v, spy = -(cos(phi) * u + sin(phi) * v);
      const double vv[4] = {spx, spy, pivot * spx, pivot * spy};
      double rd[4];
      for (int r = 0; r < 4; ++r) rd[r] = A[4 * r] * vv[0] + A[4 * r + 1] * vv[1] + A[4 * r + 2] * vv[2] + A[4 * r + 3] * vv[3];
      for (int c = 0; c < n_cam; ++c) {
        const double L = P[18 + c], a = defocus + L;
        const double bx = vv[2] * defocus + (rd[0] + off[0]) + (rd[2] + off[2]) * L;
        const double by = vv[3] * defocus + (rd[1] + off[1]) + (rd[3] + off[3]) * L;
        const float* t = T + ((int64_t)c * S + s) * J * 2;
        double M[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int64_t q = 0; q < J; ++q) {
          const double tx = th[q % n_side], ty = th[q / n_side];
          const double rx = tx * a + bx - (double)t[2 * q], ry = ty * a + by - (double)t[2 * q + 1];
          M[0] += rx; M[1] += rx * tx; M[2] += rx * ty;
          M[3] += ry; M[4] += ry * tx; M[5] += ry * ty;
          M[6] += rx * rx + ry * ry;
        }
        for (int k = 0; k < 16; ++k) {
          const int r = k >> 2, col = k & 3;
          g[k] += 2.0 * vv[col] * ((r >= 2) ? L : 1.0) * ((r == 0 || r == 2) ? M[0] : M[3]);
        }
        const double kap = pivot, d = defocus;
        const double dbx_dx = kap * d + (A[0] + A[2] * kap) + (A[8] + A[10] * kap) * L;
        const double dbx_dy = (A[1] + A[3] * kap) + (A[9] + A[11] * kap) * L;
        const double dby_dx = (A[4] + A[6] * kap) + (A[12] + A[14] * kap) * L;
        const double dby_dy = kap * d + (A[5] + A[7] * kap) + (A[13] + A[15] * kap) * L;
        const double u1 = (double)i - 0.5 * (scan_h - 1), v1 = (double)j - 0.5 * (scan_w - 1);
        const double sx1 = -sin(phi) * u1 + cos(phi) * v1, sy1 = -(cos(phi) * u1 + sin(phi) * v1);
        const double sx2 = spy * (pi / 180.0), sy2 = -spx * (pi / 180.0);
        g[16] += 2.0 * ((dbx_dx * sx1 + dbx_dy * sy1) * M[0] + (dby_dx * sx1 + dby_dy * sy1) * M[3]);
        g[17] += 2.0 * ((dbx_dx * sx2 + dbx_dy * sy2) * M[0] + (dby_dx * sx2 + dby_dy * sy2) * M[3]);
        g[18 + c] += 2.0 * (M[1] + (rd[2] + off[2]) * M[0] + M[5] + (rd[3] + off[3]) * M[3]);
        g[np - 1] += M[6];
      }
    }
#pragma omp critical
    for (int k = 0; k < np; ++k) out[k] += g[k];
    free(g);
  }
  free(th);
}
