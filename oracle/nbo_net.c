/*
 * nbo_net.c -- CPU oracle of the learned element preconditioner.
 *
 * TEST INFRASTRUCTURE ONLY (see nbo.h).  No reference counterpart: the
 * network is a new feature; this file restates the library's forward pass,
 * loss and parameter gradient in the library's fixed evaluation order.
 *
 * Pinning: tests/test_oracle_net.py checks the forward pass against an
 * independent numpy float64 implementation and the analytic gradient against
 * central finite differences of that implementation; with zero output
 * weights the preconditioner must equal Jacobi exactly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "nbo.h"

enum { P = 512, C = 32, G = 7, TAPS = 27 };
enum { OWR = 0, OWG = 864, OB1 = 1088, OW2 = 1120, OB2 = 2144, OW3 = 2176, OB3 = 3040, NPAR = 3041 };

static int tap_nb(int n, int tap)
{
  const int i = n & 7, j = (n >> 3) & 7, k = n >> 6;
  const int ii = i + tap % 3 - 1, jj = j + (tap / 3) % 3 - 1, kk = k + tap / 9 - 1;
  if (ii < 0 || ii > 7 || jj < 0 || jj > 7 || kk < 0 || kk > 7) return -1;
  return ii + 8 * jj + 64 * kk;
}

typedef struct {
  double s;
  float rh[P];
  float g[P][G];
  float h1[P][C], h2[P][C];
  float T[P][TAPS];
  float y[P];
} elem_fwd;

/* forward of one element: r (double, 512), feat [7][512] (float) */
static void elem_forward(const float* W, const double* r, const float* feat, elem_fwd* a)
{
  double s = 0.0;
  for (int n = 0; n < P; ++n) s = fmax(s, fabs(r[n]));
  a->s = s;
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  for (int n = 0; n < P; ++n) {
    a->rh[n] = (float)(r[n] * inv);
    for (int q = 0; q < G; ++q) a->g[n][q] = feat[q * P + n];
  }
  for (int n = 0; n < P; ++n) {
    float h1[C], h2[C];
    for (int c = 0; c < C; ++c) h1[c] = W[OB1 + c];
    for (int q = 0; q < G; ++q)
      for (int c = 0; c < C; ++c) h1[c] = fmaf(W[OWG + q * C + c], a->g[n][q], h1[c]);
    for (int tap = 0; tap < TAPS; ++tap) {
      const int nb = tap_nb(n, tap);
      if (nb < 0) continue;
      const float v = a->rh[nb];
      for (int c = 0; c < C; ++c) h1[c] = fmaf(W[OWR + tap * C + c], v, h1[c]);
    }
    for (int c = 0; c < C; ++c) h1[c] = h1[c] > 0.f ? h1[c] : 0.f;
    for (int c = 0; c < C; ++c) h2[c] = W[OB2 + c];
    for (int ci = 0; ci < C; ++ci)
      for (int c = 0; c < C; ++c) h2[c] = fmaf(W[OW2 + ci * C + c], h1[ci], h2[c]);
    for (int c = 0; c < C; ++c) h2[c] = h2[c] > 0.f ? h2[c] : 0.f;
    for (int tap = 0; tap < TAPS; ++tap) {
      float t = 0.f;
      for (int c = 0; c < C; ++c) t = fmaf(W[OW3 + tap * C + c], h2[c], t);
      a->T[n][tap] = t;
    }
    memcpy(a->h1[n], h1, sizeof(h1));
    memcpy(a->h2[n], h2, sizeof(h2));
  }
  for (int n = 0; n < P; ++n) {
    float y = W[OB3];
    for (int tap = 0; tap < TAPS; ++tap) {
      const int nb = tap_nb(n, tap);
      if (nb >= 0) y = y + a->T[nb][tap];
    }
    a->y[n] = y;
  }
}

/* zl = w (invDiag (r + s y)) for every element (w may be NULL: weight 1) */
void nbo_net_apply(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                   const double* w, double* zl)
{
#pragma omp parallel
  {
    elem_fwd* a = (elem_fwd*)malloc(sizeof(elem_fwd));
#pragma omp for schedule(static)
    for (int32_t e = 0; e < E; ++e) {
      elem_forward(W, r + (int64_t)e * P, feat + (int64_t)e * G * P, a);
      for (int n = 0; n < P; ++n) {
        const int64_t id = (int64_t)e * P + n;
        const double wn = w ? w[id] : 1.0;
        zl[id] = wn * (invDiag[id] * (r[id] + a->s * (double)a->y[n]));
      }
    }
    free(a);
  }
}

/* Loss and gradient over samples (pair, element), chunk samples per
 * workgroup in the library: lane-owned accumulators per chunk, chunks summed
 * in order.  R, Ev: float pairs with stride pairStride; grad[NPAR]. */
void nbo_net_gradient(int32_t nsamples, int32_t chunk, const int32_t* samples, int64_t pairStride,
                      const float* R, const float* Ev, const double* invNorm, const float* W, const float* feat,
                      const double* invDiag, const double* Gz, float* grad, double* loss_out)
{
  const int nblocks = (nsamples + chunk - 1) / chunk;
  float* part = (float*)calloc((size_t)nblocks * NPAR, sizeof(float));
  double* lpart = (double*)calloc(nblocks, sizeof(double));
#pragma omp parallel
  {
    elem_fwd* a = (elem_fwd*)malloc(sizeof(elem_fwd));
    double* rr = (double*)malloc(sizeof(double) * P);
    float* dy = (float*)malloc(sizeof(float) * P);
    float (*dT)[TAPS] = malloc(sizeof(float) * P * TAPS);
    float (*dp2)[C] = malloc(sizeof(float) * P * C);
    float (*dp1)[C] = malloc(sizeof(float) * P * C);
#pragma omp for schedule(static)
    for (int b = 0; b < nblocks; ++b) {
      float* gp = part + (size_t)b * NPAR;
      double loss = 0.0;
      const int s0 = b * chunk, s1 = (s0 + chunk < nsamples) ? s0 + chunk : nsamples;
      for (int si = s0; si < s1; ++si) {
        const int p = samples[2 * si], e = samples[2 * si + 1];
        const float* Rp = R + (int64_t)p * pairStride + (int64_t)e * P;
        const float* Ep = Ev + (int64_t)p * pairStride + (int64_t)e * P;
        for (int n = 0; n < P; ++n) rr[n] = (double)Rp[n];
        elem_forward(W, rr, feat + (int64_t)e * G * P, a);
        const double wn = Gz ? 0.0 : invNorm[p];
        double lane[P];
        for (int n = 0; n < P; ++n) {
          const double di = invDiag[(int64_t)e * P + n];
          if (Gz) {  /* external loss: dL/du given, u = invDiag (r + s y) */
            dy[n] = (float)(Gz[(int64_t)e * P + n] * di * a->s);
            continue;
          }
          const double z = di * (rr[n] + a->s * (double)a->y[n]);
          const double res = z - (double)Ep[n];
          lane[n] = res * res * wn;
          dy[n] = (float)(2.0 * res * wn * di * a->s);
        }
        if (!Gz) { /* per-wave halving tree, waves added in order */
          double t = 0.0;
          for (int w = 0; w < P / 64; ++w) {
            double v[64];
            for (int l = 0; l < 64; ++l) v[l] = lane[w * 64 + l];
            for (int h = 32; h >= 1; h >>= 1)
              for (int l = 0; l < h; ++l) v[l] = v[l] + v[l + h];
            t = (w == 0) ? v[0] : t + v[0];
          }
          loss = loss + t;
        }
        for (int n = 0; n < P; ++n) {
          for (int tap = 0; tap < TAPS; ++tap) {
            const int m = tap_nb(n, 26 - tap);
            dT[n][tap] = m >= 0 ? dy[m] : 0.f;
          }
          for (int c = 0; c < C; ++c) {
            float acc = 0.f;
            for (int tap = 0; tap < TAPS; ++tap) acc = fmaf(W[OW3 + tap * C + c], dT[n][tap], acc);
            dp2[n][c] = a->h2[n][c] > 0.f ? acc : 0.f;
          }
          for (int ci = 0; ci < C; ++ci) {
            float acc = 0.f;
            for (int c = 0; c < C; ++c) acc = fmaf(W[OW2 + ci * C + c], dp2[n][c], acc);
            dp1[n][ci] = a->h1[n][ci] > 0.f ? acc : 0.f;
          }
        }
        for (int q = 0; q < TAPS * C; ++q) {  /* dW3 */
          const int tap = q / C, c = q % C;
          float acc = 0.f;
          for (int m = 0; m < P; ++m) acc = fmaf(a->h2[m][c], dT[m][tap], acc);
          gp[OW3 + q] = gp[OW3 + q] + acc;
        }
        {
          float acc = 0.f;
          for (int m = 0; m < P; ++m) acc = acc + dy[m];
          gp[OB3] = gp[OB3] + acc;
        }
        for (int q = 0; q < C * C; ++q) {  /* dW2 */
          const int ci = q / C, c = q % C;
          float acc = 0.f;
          for (int m = 0; m < P; ++m) acc = fmaf(a->h1[m][ci], dp2[m][c], acc);
          gp[OW2 + q] = gp[OW2 + q] + acc;
        }
        for (int c = 0; c < C; ++c) {
          float acc = 0.f;
          for (int m = 0; m < P; ++m) acc = acc + dp2[m][c];
          gp[OB2 + c] = gp[OB2 + c] + acc;
        }
        for (int q = 0; q < TAPS * C; ++q) {  /* dWr */
          const int tap = q / C, c = q % C;
          float acc = 0.f;
          for (int m = 0; m < P; ++m) {
            const int nb = tap_nb(m, tap);
            if (nb >= 0) acc = fmaf(dp1[m][c], a->rh[nb], acc);
          }
          gp[OWR + q] = gp[OWR + q] + acc;
        }
        for (int q = 0; q < G * C; ++q) {  /* dWg */
          const int g = q / C, c = q % C;
          float acc = 0.f;
          for (int m = 0; m < P; ++m) acc = fmaf(dp1[m][c], a->g[m][g], acc);
          gp[OWG + q] = gp[OWG + q] + acc;
        }
        for (int c = 0; c < C; ++c) {
          float acc = 0.f;
          for (int m = 0; m < P; ++m) acc = acc + dp1[m][c];
          gp[OB1 + c] = gp[OB1 + c] + acc;
        }
      }
      lpart[b] = loss;
    }
    free(a); free(rr); free(dy); free(dT); free(dp2); free(dp1);
  }
  for (int q = 0; q < NPAR; ++q) {
    float acc = 0.f;
    for (int b = 0; b < nblocks; ++b) acc = acc + part[(size_t)b * NPAR + q];
    grad[q] = acc;
  }
  double l = 0.0;
  for (int b = 0; b < nblocks; ++b) l = l + lpart[b];
  *loss_out = l;
  free(part); free(lpart);
}

/* Features of a mesh: each ggeo factor divided by its RMS over the mesh.  The
 * sums of squares follow the library's two-phase order: block b takes
 * elements [16 b, 16 b + 16), lane t of 256 runs an fma chain per factor over
 * those elements and nodes t, t+256, ...; the 256-lane tree per block; the
 * block partials combined by the partials tree. */
void nbo_net_features(int32_t E, int Np, const double* ggeo, float* feat)
{
  const int64_t nb = ((int64_t)E + 15) / 16;
  double* parts = (double*)malloc(sizeof(double) * G * (size_t)nb);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    double lane[G][256];
    for (int t = 0; t < 256; ++t) {
      double acc[G] = {0, 0, 0, 0, 0, 0, 0};
      const int64_t e1 = (b + 1) * 16 < E ? (b + 1) * 16 : E;
      for (int64_t e = b * 16; e < e1; ++e)
        for (int n = t; n < Np; n += 256)
          for (int q = 0; q < G; ++q) {
            const double g = ggeo[(e * G + q) * Np + n];
            acc[q] = fma(g, g, acc[q]);
          }
      for (int q = 0; q < G; ++q) lane[q][t] = acc[q];
    }
    for (int q = 0; q < G; ++q) parts[q * nb + b] = nbo_tree256(lane[q]);
  }
  double inv[G];
  for (int q = 0; q < G; ++q) {
    const double rms = sqrt(nbo_tree_sum_partials(nb, parts + q * nb) / ((double)E * Np));
    inv[q] = rms > 0 ? 1.0 / rms : 0.0;
  }
  free(parts);
  for (int64_t i = 0; i < (int64_t)E * G * Np; ++i) feat[i] = (float)(ggeo[i] * inv[(i / Np) % G]);
}

/* 1 / sum of squares of a float vector in the library's order (lane t of 256
 * runs an fma chain over i = t, t+256, ...; the 256-lane tree) */
double nbo_sumsq_inv_f(int64_t n, const float* a)
{
  double lane[256];
  for (int t = 0; t < 256; ++t) {
    double acc = 0.0;
    for (int64_t i = t; i < n; i += 256) acc = fma((double)a[i], (double)a[i], acc);
    lane[t] = acc;
  }
  return 1.0 / nbo_tree256(lane);
}

/* Matrix-core inference (library: k_net_apply_mma, nbx_net_mma.hip): the
 * same network with the layer inputs (features, hidden activations) and the
 * weights rounded to bf16 (nearest even), products accumulated in double and
 * rounded to float (the products are exact in FP32; the matrix cores'
 * summation order is not specified, so agreement is to FP32 rounding). */
static float bfr(float x) { return nbo_bf16r(x); }

void nbo_net_apply_mma(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                       const double* w, double* zl)
{
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < E; ++e) {
    const double* re = r + (int64_t)e * P;
    double s = 0.0;
    for (int n = 0; n < P; ++n) s = fmax(s, fabs(re[n]));
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    float rh[P];
    for (int n = 0; n < P; ++n) rh[n] = (float)(re[n] * inv);
    static float T[P][TAPS];
#pragma omp threadprivate(T)
    for (int n = 0; n < P; ++n) {
      float fv[34];
      for (int q = 0; q < G; ++q) fv[q] = bfr(feat[((int64_t)e * G + q) * P + n]);
      for (int tap = 0; tap < TAPS; ++tap) {
        const int nb = tap_nb(n, tap);
        fv[7 + tap] = bfr(nb >= 0 ? rh[nb] : 0.f);
      }
      float h1[C], h2[C];
      for (int c = 0; c < C; ++c) {
        double acc = W[OB1 + c];
        for (int f = 0; f < 7; ++f) acc += (double)bfr(W[OWG + f * C + c]) * fv[f];
        for (int tap = 0; tap < TAPS; ++tap) acc += (double)bfr(W[OWR + tap * C + c]) * fv[7 + tap];
        const float v = (float)acc;
        h1[c] = bfr(v > 0.f ? v : 0.f);
      }
      for (int c = 0; c < C; ++c) {
        double acc = W[OB2 + c];
        for (int ci = 0; ci < C; ++ci) acc += (double)bfr(W[OW2 + ci * C + c]) * h1[ci];
        const float v = (float)acc;
        h2[c] = bfr(v > 0.f ? v : 0.f);
      }
      for (int tap = 0; tap < TAPS; ++tap) {
        double acc = 0.0;
        for (int c = 0; c < C; ++c) acc += (double)bfr(W[OW3 + tap * C + c]) * h2[c];
        T[n][tap] = (float)acc;
      }
    }
    for (int n = 0; n < P; ++n) {
      float y = W[OB3];
      for (int tap = 0; tap < TAPS; ++tap) {
        const int nb = tap_nb(n, tap);
        if (nb >= 0) y = y + T[nb][tap];
      }
      const int64_t id = (int64_t)e * P + n;
      const double wn = w ? w[id] : 1.0;
      zl[id] = wn * (invDiag[id] * (re[n] + s * (double)y));
    }
  }
}
