/*
 * nbo_pcg.c -- CPU oracle of nekBone's BP operator and PCG loop.
 *
 * TEST INFRASTRUCTURE ONLY (see nbo.h).
 *
 * Restates BPPCG (nekBone/solve.cpp:44-113), BPAxOperator (:181-235),
 * updateJacobi (:146-165), BPWeightedInnerProduct[2] (:237-287),
 * BPScaledAdd (:289-293) and BPUpdatePCG (:136-142, utils.okl:795-807) in the
 * library's arithmetic order: one fma per element-wise update, dots in the
 * fixed tile tree, scalars evaluated as C expressions in the reference's form.
 *
 * fused != 0 restates the library's fused schedule: pAp = sum_local p (A_L p)
 * taken from the operator's local output before the gather-scatter (equal to
 * the weighted dot in real arithmetic for continuous, masked p).
 *
 * Pinning: the operator is pinned through nbo_axhelm / nbo_gs_many; the loop
 * by CG invariants on small problems (orthogonality of residuals, A-conjugacy
 * of directions, monotone A-norm error), by the dense direct solve of the
 * assembled system, and by the manufactured cosine solution of
 * nekBone/main.cpp:105-132 (tests/test_oracle_pcg.py).
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include "nbo.h"

#define TILE 2048

/* partial sums of NV values per tile in the library's tree, then combined */
#define TILE_TREE(NVV, N_, OUT, BODY)                                              \
  do {                                                                             \
    const int64_t nb_ = ((N_) + TILE - 1) / TILE;                                  \
    double* parts_ = (double*)malloc(sizeof(double) * (NVV) * (nb_ > 0 ? nb_ : 1)); \
    _Pragma("omp parallel for schedule(static)")                                   \
    for (int64_t b = 0; b < nb_; ++b) {                                            \
      double lane_[NVV][256];                                                      \
      for (int t = 0; t < 256; ++t) {                                              \
        double acc[NVV];                                                           \
        for (int c = 0; c < (NVV); ++c) acc[c] = 0.0;                              \
        for (int u = 0; u < 8; ++u) {                                              \
          const int64_t id = b * TILE + u * 256 + t;                               \
          if (id < (N_)) { BODY }                                                  \
        }                                                                          \
        for (int c = 0; c < (NVV); ++c) lane_[c][t] = acc[c];                      \
      }                                                                            \
      for (int c = 0; c < (NVV); ++c) parts_[c * nb_ + b] = nbo_tree256(lane_[c]); \
    }                                                                              \
    for (int c = 0; c < (NVV); ++c) (OUT)[c] = nbo_tree_sum_partials(nb_, parts_ + c * nb_); \
    free(parts_);                                                                  \
  } while (0)

void nbo_bp_apply(const nbo_bp* bp, const double* q, double* Aq)
{
  nbo_axhelm(bp->N, bp->Nf, bp->bk, bp->E, bp->offset, NULL, bp->ggeo, bp->D, bp->lambda, q, Aq);
  nbo_gs_many(bp->ogs, Aq, bp->Nf, bp->offset, NBO_DOUBLE, NBO_ADD);
  for (int f = 0; f < bp->Nf; ++f)
    for (int32_t m = 0; m < bp->Nmasked; ++m) Aq[bp->maskIds[m] + (int64_t)f * bp->offset] = 0.0;
}

double nbo_bp_wdot(const nbo_bp* bp, const double* x, const double* y)
{
  double out[1];
  const int64_t N = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  TILE_TREE(1, N, out, {
    const double wid = bp->w[id];
    for (int f = 0; f < bp->Nf; ++f) {
      const int64_t iid = id + (int64_t)f * bp->offset;
      acc[0] = fma(wid * x[iid], y[iid], acc[0]);
    }
  });
  return out[0];
}

/* Element shares of pAp = sum_local p (A_L p) in the library's fused order:
 * lane (i,j) runs an fma chain over fields then k -- for three fields of the
 * non-BK operator with 8 <= Nq <= 12 and FP64/FP32 geometry (one thread group
 * per field) a chain over k per field and then
 * (field0 + field1) + field2; the Nq*Nq lanes are summed by a halving tree
 * over the next power of two (zero padded); the element shares are combined
 * like tile partials. */
static double fused_pap(const nbo_bp* bp, const double* p, const double* Ap, int bf16_geometry)
{
  const int Nq = bp->N + 1, Nq2 = Nq * Nq, Np = Nq2 * Nq;
  int P2 = 1;
  while (P2 < Nq2) P2 *= 2;
  double* part = (double*)malloc(sizeof(double) * (bp->E > 0 ? bp->E : 1));
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < bp->E; ++e) {
    double lanes[256];
    for (int t = 0; t < P2; ++t) lanes[t] = 0.0;
    const int field_groups = bp->Nf == 3 && !bp->bk && Nq2 >= 64 && Nq <= 12 && !bf16_geometry;
    for (int t = 0; t < Nq2; ++t) {
      double acc = 0.0, accf[3] = {0.0, 0.0, 0.0};
      for (int f = 0; f < bp->Nf; ++f)
        for (int k = 0; k < Nq; ++k) {
          const int64_t n = (int64_t)e * Np + (int64_t)k * Nq2 + t + (int64_t)f * bp->offset;
          if (field_groups)
            accf[f] = fma(p[n], Ap[n], accf[f]);
          else
            acc = fma(p[n], Ap[n], acc);
        }
      lanes[t] = field_groups ? (accf[0] + accf[1]) + accf[2] : acc;
    }
    for (int h = P2 / 2; h >= 1; h >>= 1)
      for (int t = 0; t < h; ++t) lanes[t] = lanes[t] + lanes[t + h];
    part[e] = lanes[0];
  }
  const double r = nbo_tree_sum_partials(bp->E, part);
  free(part);
  return r;
}

/* operands of a low-precision operator: float copies (prec 1) or bf16
 * values held in float (prec 2, round-to-nearest-even of the float) */
typedef struct { float *g, *D, *lam; int bf16; } lo_ops;

static float to_bf16f(double v) { return nbo_bf16r((float)v); }

static void lo_make(const nbo_bp* bp, int prec, lo_ops* lo)
{
  const int Nq = bp->N + 1, Np = Nq * Nq * Nq;
  const int64_t ng = (int64_t)bp->E * 7 * Np, nl = 2 * (int64_t)bp->offset;
  lo->bf16 = prec == 2;
  lo->g = (float*)malloc(sizeof(float) * ng);
  lo->lam = (float*)malloc(sizeof(float) * nl);
  lo->D = (float*)malloc(sizeof(float) * Nq * Nq);
  for (int i = 0; i < Nq * Nq; ++i) lo->D[i] = (float)bp->D[i];
  for (int64_t i = 0; i < ng; ++i) lo->g[i] = prec == 2 ? to_bf16f(bp->ggeo[i]) : (float)bp->ggeo[i];
  for (int64_t i = 0; i < nl; ++i) lo->lam[i] = prec == 2 ? to_bf16f(bp->lambda[i]) : (float)bp->lambda[i];
}

static void lo_free(lo_ops* lo) { free(lo->g); free(lo->lam); free(lo->D); }

/* Ap (FP64) = A_lo p: p rounded to float, the float operator, result widened */
static void lo_axhelm(const nbo_bp* bp, const lo_ops* lo, const double* p, double* Ap)
{
  const int64_t Nall = (int64_t)bp->Nf * bp->offset;
  float* p32 = (float*)malloc(sizeof(float) * Nall);
  float* a32 = (float*)malloc(sizeof(float) * Nall);
  #pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < Nall; ++n) p32[n] = (float)p[n];
  nbo_axhelm_f32(bp->N, bp->Nf, bp->bk, bp->E, bp->offset, lo->g, lo->D, lo->lam, p32, a32);
  #pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < Nall; ++n) Ap[n] = (double)a32[n];
  free(p32); free(a32);
}

static void local_jacobi(const nbo_bp* bp, double* invDiag)
{
  const int64_t N = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  nbo_jacobi_local(bp->N, bp->E, bp->offset, NULL, bp->ggeo, bp->D, bp->lambda, invDiag);
  nbo_gs_many(bp->ogs, invDiag, 1, bp->offset, NBO_DOUBLE, NBO_ADD);
  #pragma omp parallel for schedule(static)
  for (int64_t n = 0; n < N; ++n) invDiag[n] = 1.0 / invDiag[n];
}

typedef struct {
  const float* W;      /* network weights (bf16 values) */
  const float* feat;   /* normalised geometric factors [E][7][512] */
  int window;          /* stall test: rdotr(it) > ratio * rdotr(it - window) -> Jacobi */
  double ratio;
  const nbo_mg* mg;    /* non-NULL: the preconditioner is one multigrid V-cycle instead */
} learned_ctx;

/* PCG core; lo == NULL: FP64 operator.  initial: r -= A x first.
 * ln != NULL: learned preconditioner (flexible beta, stall fallback). */
static int pcg_core_ext(const nbo_bp* bp, int jacobi, int flexible, int fused, const lo_ops* lo,
                        const double* invDiag, const learned_ctx* ln, double* r, double* x, double tol,
                        double abs_floor, int maxit, int initial, double* hist_rdotr, double* hist_alpha,
                        nbo_pcg_stats* st)
{
  const int Np = (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  const int64_t N = (int64_t)bp->E * Np;
  const int64_t Nall = (int64_t)bp->Nf * bp->offset;
  const int fixed = tol > 0 ? 0 : 1;
  double* p = (double*)calloc(Nall, sizeof(double));
  double* z = (double*)calloc(Nall, sizeof(double));
  double* Ap = (double*)calloc(Nall, sizeof(double));
  double* Ax = (double*)calloc(Nall, sizeof(double));
  double rdotz1 = 1.0, rdotz2 = 0.0, alpha = 0.0, beta = 0.0, TOL = 0.0, rdotr = 0.0, rdotr0 = 0.0;
  int converged = 0, breakdown = 0, iterations = maxit, it, fallback = 0, fallback_iter = 0;
  double* hist_own = NULL;
  if (ln) {
    flexible = 1;
    if (!hist_rdotr) hist_rdotr = hist_own = (double*)calloc((size_t)maxit + 2, sizeof(double));
  }

  if (initial) {  /* Ax = A x; r = r - Ax (solve.cpp:68-69) */
    nbo_bp_apply(bp, x, Ax);
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < Nall; ++n) r[n] = -1.0 * Ax[n] + 1.0 * r[n];
  }
  for (it = 1;; ++it) {
    if (it > maxit) { iterations = maxit; break; }
    /* z = M r; (r.z)_w, (r.r)_w [, (z.Ap)_w]  (solve.cpp:77-83, 94) */
    double d[3];
    const double* zsrc = jacobi ? z : r;
    if (ln && ln->mg) {  /* z = one V-cycle applied to r */
      nbo_mg_vcycle(ln->mg, 0, r, z);
      zsrc = z;
    } else if (ln) {  /* z = gs(w invDiag (r + s f(r))), Jacobi after a stall */
      if (fallback) {
        #pragma omp parallel for schedule(static)
        for (int64_t n = 0; n < N; ++n) z[n] = bp->w[n] * (invDiag[n] * r[n]);
      } else {
        nbo_net_apply(bp->E, ln->W, ln->feat, r, invDiag, bp->w, z);
      }
      nbo_gs_many(bp->ogs, z, bp->Nf, bp->offset, NBO_DOUBLE, NBO_ADD);
      for (int32_t m = 0; m < bp->Nmasked; ++m) z[bp->maskIds[m]] = 0.0;
      zsrc = z;
    }
    if (ln) TILE_TREE(3, N, d, {
      const double wid = bp->w[id];
      for (int f = 0; f < bp->Nf; ++f) {
        const int64_t iid = id + (int64_t)f * bp->offset;
        const double rf = r[iid];
        const double zf = z[iid];
        acc[0] = fma(wid * rf, zf, acc[0]);
        acc[1] = fma(wid * rf, rf, acc[1]);
        acc[2] = fma(wid * zf, Ap[iid], acc[2]);
      }
    });
    else TILE_TREE(3, N, d, {
      const double wid = bp->w[id];
      for (int f = 0; f < bp->Nf; ++f) {
        const int64_t iid = id + (int64_t)f * bp->offset;
        const double rf = r[iid];
        double zf = rf;
        if (jacobi) { zf = invDiag[id] * rf; z[iid] = zf; }
        acc[0] = fma(wid * rf, zf, acc[0]);
        acc[1] = fma(wid * rf, rf, acc[1]);
        if (flexible) acc[2] = fma(wid * zf, Ap[iid], acc[2]);
      }
    });
    rdotz2 = rdotz1;
    rdotz1 = d[0];
    rdotr = d[1];
    if (hist_rdotr) hist_rdotr[it] = rdotr;
    if (it == 1) { rdotr0 = rdotr; TOL = fmax(tol * tol * rdotr, abs_floor); }  /* solve.cpp:85 */
    if (!isfinite(rdotr)) { breakdown = 1; iterations = it - 1; break; }
    if (rdotr <= TOL && !fixed) { converged = 1; iterations = it - 1; break; }  /* :90 */
    if (ln && !ln->mg && !fallback && it > ln->window &&
        !(rdotr <= ln->ratio * hist_rdotr[it - ln->window])) {
      fallback = 1;
      fallback_iter = it;
    }
    if (flexible) beta = -alpha * d[2] / rdotz2;                                    /* :92-95 */
    else beta = (it == 1) ? 0.0 : rdotz1 / rdotz2;                                 /* :97 */
    /* p = z + beta p (:100) */
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < Nall; ++n) p[n] = fma(beta, p[n], zsrc[n]);
    /* Ap = A p; pAp = (p.Ap)_w; alpha = rdotz/pAp (:102-104) */
    double pAp;
    if (fused || lo || ln) {  /* pAp from the local operator output, then gs + mask */
      if (lo) lo_axhelm(bp, lo, p, Ap);
      else nbo_axhelm(bp->N, bp->Nf, bp->bk, bp->E, bp->offset, NULL, bp->ggeo, bp->D, bp->lambda, p, Ap);
      pAp = fused_pap(bp, p, Ap, lo && lo->bf16);
      nbo_gs_many(bp->ogs, Ap, bp->Nf, bp->offset, NBO_DOUBLE, NBO_ADD);
      for (int f = 0; f < bp->Nf; ++f)
        for (int32_t m = 0; m < bp->Nmasked; ++m) Ap[bp->maskIds[m] + (int64_t)f * bp->offset] = 0.0;
    } else {
      nbo_bp_apply(bp, p, Ap);
      pAp = nbo_bp_wdot(bp, p, Ap);
    }
    alpha = rdotz1 / pAp;
    if (hist_alpha) hist_alpha[it] = alpha;
    if (!isfinite(alpha)) { breakdown = 1; iterations = it - 1; break; }
    /* x += alpha p; r -= alpha Ap (:108) */
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < Nall; ++n) {
      x[n] = fma(alpha, p[n], x[n]);
      r[n] = fma(-alpha, Ap[n], r[n]);
    }
  }
  if (st) {
    st->iterations = iterations; st->converged = converged; st->breakdown = breakdown; st->iter = it;
    st->rdotr0 = rdotr0; st->rdotr = rdotr; st->TOL = TOL;
    st->fallback_iter = fallback ? fallback_iter : 0;
  }
  free(p); free(z); free(Ap); free(Ax); free(hist_own);
  return iterations;
}

static int pcg_core(const nbo_bp* bp, int jacobi, int flexible, int fused, const lo_ops* lo, const double* invDiag,
                    double* r, double* x, double tol, double abs_floor, int maxit, int initial,
                    double* hist_rdotr, double* hist_alpha, nbo_pcg_stats* st)
{
  return pcg_core_ext(bp, jacobi, flexible, fused, lo, invDiag, NULL, r, x, tol, abs_floor, maxit, initial,
                      hist_rdotr, hist_alpha, st);
}

int nbo_bp_pcg(const nbo_bp* bp, int jacobi, int flexible, int fused, double* r, double* x, double tol, int maxit,
               double* hist_rdotr, double* hist_alpha, nbo_pcg_stats* st)
{
  const int64_t N = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  double* invDiag = NULL;
  if (jacobi) {
    /* the initial residual is formed before the Jacobi setup (solve.cpp:68-72);
     * the diagonal does not depend on r, so the order does not matter */
    invDiag = (double*)malloc(sizeof(double) * N);
    local_jacobi(bp, invDiag);
  }
  const int its = pcg_core(bp, jacobi, flexible, fused, NULL, invDiag, r, x, tol, tol * tol, maxit, 1,
                           hist_rdotr, hist_alpha, st);
  free(invDiag);
  return its;
}

static int pcg_mixed_impl(const nbo_bp* bp, int jacobi, int flexible, int prec, const learned_ctx* ln, double* r,
                          double* x, double tol, int maxit, double inner_tol, int max_cycles, nbo_pcg_stats* st);

/* Iterative refinement around PCG with a low-precision operator (library's
 * nbx_bp_pcg_mixed): r = b - A64 x; each cycle solves A_lo d = r from d = 0
 * to inner_tol (no absolute floor), x += d; stop on the true residual. */
int nbo_bp_pcg_mixed(const nbo_bp* bp, int jacobi, int flexible, int prec, double* r, double* x, double tol,
                     int maxit, double inner_tol, int max_cycles, nbo_pcg_stats* st)
{
  return pcg_mixed_impl(bp, jacobi, flexible, prec, NULL, r, x, tol, maxit, inner_tol, max_cycles, st);
}

/* The same refinement with the learned preconditioner in the inner solves
 * (library: nbx_bp_pcg_mixed on a handle with precon LEARNED): each inner
 * solve is the learned flexible PCG of nbo_bp_pcg_learned with the
 * low-precision operator, its stall test starting afresh. */
int nbo_bp_pcg_mixed_learned(const nbo_bp* bp, const float* W, const float* feat, int window, double ratio, int prec,
                             double* r, double* x, double tol, int maxit, double inner_tol, int max_cycles,
                             nbo_pcg_stats* st)
{
  const learned_ctx ln = {W, feat, window, ratio, NULL};
  return pcg_mixed_impl(bp, 1, 1, prec, &ln, r, x, tol, maxit, inner_tol, max_cycles, st);
}

static int pcg_mixed_impl(const nbo_bp* bp, int jacobi, int flexible, int prec, const learned_ctx* ln, double* r,
                          double* x, double tol, int maxit, double inner_tol, int max_cycles, nbo_pcg_stats* st)
{
  const int64_t N = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  const int64_t Nall = (int64_t)bp->Nf * bp->offset;
  lo_ops lo = {0};
  if (prec != 0) lo_make(bp, prec, &lo);
  double* b = (double*)malloc(sizeof(double) * Nall);
  double* Ax = (double*)malloc(sizeof(double) * Nall);
  double* d = (double*)malloc(sizeof(double) * Nall);
  double* invDiag = NULL;
  memcpy(b, r, sizeof(double) * Nall);
  int total = 0, cycles = 0, conv = 0, breakdown = 0;
  double rr0 = 0.0, rr = 0.0;
  for (;;) {
    nbo_bp_apply(bp, x, Ax);
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < Nall; ++n) r[n] = -1.0 * Ax[n] + 1.0 * b[n];
    rr = nbo_bp_wdot(bp, r, r);
    if (cycles == 0) rr0 = rr;
    if (rr <= tol * tol * rr0) { conv = 1; break; }
    if (cycles >= max_cycles || total >= maxit || breakdown) break;
    if (jacobi && !invDiag) { invDiag = (double*)malloc(sizeof(double) * N); local_jacobi(bp, invDiag); }
    memset(d, 0, sizeof(double) * Nall);
    nbo_pcg_stats s;
    pcg_core_ext(bp, jacobi, flexible, 1, prec ? &lo : NULL, invDiag, ln, r, d, inner_tol, 0.0, maxit - total, 0,
                 NULL, NULL, &s);
    total += s.iterations;
    breakdown = s.breakdown;
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < Nall; ++n) x[n] = x[n] + d[n];
    ++cycles;
    if (s.iterations == 0) break;
  }
  if (st) {
    st->iterations = total; st->converged = conv; st->breakdown = breakdown; st->iter = cycles;
    st->rdotr0 = rr0; st->rdotr = rr; st->TOL = tol * tol * rr0;
  }
  free(b); free(Ax); free(d); free(invDiag);
  if (prec != 0) lo_free(&lo);
  return total;
}

/* BPPCG with the learned preconditioner (library: nbx_bp_pcg, precon LEARNED) */
int nbo_bp_pcg_learned(const nbo_bp* bp, const float* W, const float* feat, int window, double ratio, double* r,
                       double* x, double tol, int maxit, double* hist_rdotr, double* hist_alpha, nbo_pcg_stats* st)
{
  const int64_t N = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  double* invDiag = (double*)malloc(sizeof(double) * N);
  local_jacobi(bp, invDiag);
  const learned_ctx ln = {W, feat, window, ratio, NULL};
  const int its = pcg_core_ext(bp, 1, 1, 1, NULL, invDiag, &ln, r, x, tol, tol * tol, maxit, 1, hist_rdotr,
                               hist_alpha, st);
  free(invDiag);
  return its;
}

/* BPZeroMean (nekBone/solve.cpp:115-134): q -= (unweighted sum of the
 * E-vector over its fields) / (E Nf Np) on the whole Nf*offset range, the sum
 * in the library's tile tree */
void nbo_bp_zero_mean(const nbo_bp* bp, double* q)
{
  const int Np = (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  const int64_t N = (int64_t)bp->E * Np;
  double d[1];
  TILE_TREE(1, N, d, {
    for (int f = 0; f < bp->Nf; ++f) acc[0] = acc[0] + q[id + (int64_t)f * bp->offset];
  });
  const double mean = d[0] / ((double)bp->E * bp->Nf * Np);
  for (int64_t i = 0; i < (int64_t)bp->Nf * bp->offset; ++i) q[i] += -mean;
}


/* ---- p-multigrid preconditioner (library: nbx_mg.hip) ------------------- */

/* J[i][a] = prod_{m != a} (zf[i] - zc[m]) / (zc[a] - zc[m]), ascending m */
void nbo_mg_interp(int Nf, int Nc, double* J)
{
  double zf[32], wf[32], zc[32], wc[32];
  nbo_gll(Nf, zf, wf);
  nbo_gll(Nc, zc, wc);
  for (int i = 0; i <= Nf; ++i)
    for (int a = 0; a <= Nc; ++a) {
      double l = 1.0;
      for (int m = 0; m <= Nc; ++m)
        if (m != a) l = l * ((zf[i] - zc[m]) / (zc[a] - zc[m]));
      J[i * (Nc + 1) + a] = l;
    }
}

void nbo_mg_cheb(double lmax, double lo, double hi, int degree, double* c1, double* c2)
{
  const double a = lo * lmax, b = hi * lmax;
  const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
  double rho = 1.0 / sigma;
  for (int s = 0; s < degree; ++s) c1[s] = c2[s] = 0.0;
  c2[0] = 1.0 / theta;
  for (int s = 1; s < degree; ++s) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    c1[s] = rn * rho;
    c2[s] = 2.0 * rn / delta;
    rho = rn;
  }
}

/* lambda_max(D^-1 A): v0 = gs(mask(U[-1,1))) normalised in the weighted
 * norm, then v <- D^-1 A v / ||D^-1 A v||_w; the last norm */
double nbo_mg_power(const nbo_bp* bp_in, const double* invD, uint64_t seed, uint64_t stream, int iters)
{
  /* field 0 only: the vector operator is block-diagonal with equal blocks */
  nbo_bp one = *bp_in;
  one.Nf = 1;
  const nbo_bp* bp = &one;
  const int64_t n = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  double* v = (double*)calloc(bp->offset, sizeof(double));
  double* w = (double*)calloc(bp->offset, sizeof(double));
  for (int64_t i = 0; i < n; ++i) v[i] = nbo_urand(seed, stream, (uint64_t)i);
  for (int32_t m = 0; m < bp->Nmasked; ++m) v[bp->maskIds[m]] = 0.0;
  nbo_gs_many(bp->ogs, v, 1, bp->offset, NBO_DOUBLE, NBO_ADD);
  double nrm = nbo_bp_wdot(bp, v, v);
  double sc = 1.0 / sqrt(nrm);
  for (int64_t i = 0; i < n; ++i) v[i] = sc * v[i];
  double lam = 0.0;
  for (int it = 0; it < iters; ++it) {
    nbo_bp_apply(bp, v, w);
    for (int64_t i = 0; i < n; ++i) w[i] = 1.0 * (invD[i] * w[i]);
    nrm = nbo_bp_wdot(bp, w, w);
    lam = sqrt(nrm);
    sc = 1.0 / lam;
    for (int64_t i = 0; i < n; ++i) v[i] = sc * w[i];
  }
  free(v);
  free(w);
  return lam;
}

static void mg_vcycle_f(const nbo_mg* mg, int l, const void* r, void* x);

/* Learned step of the fine level (library: the end of vcycle_level in
 * nbx_mg.hip): rf = the fine residual the cycle leaves (FP64 copy),
 * zl = w invDiag (rf + s f(rf)) element by element, x += mask(gs(zl)). */
static void mg_learned_step(const nbo_mg* mg, const double* rf, double* x)
{
  const nbo_bp* bp = &mg->bps[0];
  const int64_t n = (int64_t)bp->E * (bp->N + 1) * (bp->N + 1) * (bp->N + 1);
  double* zl = (double*)calloc(bp->offset, sizeof(double));
  nbo_net_apply(bp->E, mg->W, mg->feat, rf, mg->invD[0], bp->w, zl);
  nbo_gs_many(bp->ogs, zl, 1, bp->offset, NBO_DOUBLE, NBO_ADD);
  for (int32_t m = 0; m < bp->Nmasked; ++m) zl[bp->maskIds[m]] = 0.0;
  for (int64_t i = 0; i < n; ++i) x[i] = x[i] + zl[i];
  free(zl);
}

/* Coarse solve of order 1 in the library's two-pass form (nbx_mg.hip,
 * k_cg_pass_a / k_cg_pass_b): K Jacobi-PCG iterations from x = 0 with
 *   rz_k = sum over assembled nodes g (gather-scatter group order) of
 *          sum over fields fma(r_g, z_g, .); 32 nodes per partial, node j at
 *          lane 8j of the 256-lane tree (the other lanes zero), partials
 *          combined as tile partials;
 *   pAp  = element shares in the fused PCG's order (lane chains over fields
 *          then k, lane tree (a0 + a2) + (a1 + a3)); 64 shares per partial:
 *          a halving tree over each group of 16 (zero past the last element),
 *          then (S0 + S2) + (S1 + S3); partials combined as tile partials;
 *   the gathered Ap_g summed from 0 in member order, x += alpha p,
 *   r -= alpha Ap_g, z = D^-1 r at every member; a non-finite beta or alpha
 *   (the residual reached zero) is taken as 0, so the iterate stays.
 * Returns 0, or -1 when an assembled node has more than 8 members (the
 * library then runs the generic PCG, restated by pcg_core). */
static int coarse_cg(const nbo_bp* bp, const double* invD, double* r, double* x, int K)
{
  enum { EPB = 64 };
  const nbo_ogs* o = bp->ogs;
  const int32_t G = o->NlocalGather;
  for (int32_t g = 0; g < G; ++g)
    if (o->offsets[g + 1] - o->offsets[g] > 8) return -1;
  const int Nq2 = 4, Np = 8, Nf = bp->Nf;
  const int64_t n = (int64_t)bp->E * Np, Nall = (int64_t)Nf * bp->offset;
  const int64_t nbA = ((int64_t)bp->E + EPB - 1) / EPB, nbB = ((int64_t)G + 31) / 32;
  double* z = (double*)calloc((size_t)Nall, sizeof(double));
  double* p = (double*)calloc((size_t)Nall, sizeof(double));
  double* Ap = (double*)calloc((size_t)Nall, sizeof(double));
  double* share = (double*)calloc((size_t)(bp->E > 0 ? bp->E : 1), sizeof(double));
  double* pa = (double*)calloc((size_t)(nbA > 0 ? nbA : 1), sizeof(double));
  double* pb = (double*)calloc((size_t)(nbB > 0 ? nbB : 1), sizeof(double));
  double scal[2] = {0.0, 0.0};
  memset(x, 0, sizeof(double) * (size_t)Nall);
  for (int k = -1; k < K; ++k) {
    double alpha = 0.0;
    if (k >= 0) {
      /* pass A */
      const double rz = nbo_tree_sum_partials(nbB, pb);
      double beta = k == 0 ? 0.0 : rz / scal[(k - 1) & 1];
      if (!isfinite(beta)) beta = 0.0; /* r reached zero: the iterate stays */
      scal[k & 1] = rz;
      for (int f = 0; f < Nf; ++f)
        for (int64_t i = 0; i < n; ++i) {
          const int64_t id = i + (int64_t)f * bp->offset;
          p[id] = fma(beta, p[id], z[id]);
        }
      nbo_axhelm(1, Nf, bp->bk, bp->E, bp->offset, NULL, bp->ggeo, bp->D, bp->lambda, p, Ap);
#pragma omp parallel for schedule(static)
      for (int32_t e = 0; e < bp->E; ++e) {
        double lanes[4];
        for (int t = 0; t < Nq2; ++t) {
          double acc = 0.0;
          for (int f = 0; f < Nf; ++f)
            for (int kk = 0; kk < 2; ++kk) {
              const int64_t id = (int64_t)e * Np + (int64_t)kk * Nq2 + t + (int64_t)f * bp->offset;
              acc = fma(p[id], Ap[id], acc);
            }
          lanes[t] = acc;
        }
        for (int h = 2; h >= 1; h >>= 1)
          for (int t = 0; t < h; ++t) lanes[t] = lanes[t] + lanes[t + h];
        share[e] = lanes[0];
      }
      for (int64_t b = 0; b < nbA; ++b) {
        double S[4];
        for (int w = 0; w < 4; ++w) {
          double sh[16];
          for (int t = 0; t < 16; ++t) {
            const int64_t e = b * EPB + w * 16 + t;
            sh[t] = e < bp->E ? share[e] : 0.0;
          }
          for (int h = 8; h >= 1; h >>= 1)
            for (int t = 0; t < h; ++t) sh[t] = sh[t] + sh[t + h];
          S[w] = sh[0];
        }
        pa[b] = (S[0] + S[2]) + (S[1] + S[3]);
      }
      alpha = scal[k & 1] / nbo_tree_sum_partials(nbA, pa);
      if (!isfinite(alpha)) alpha = 0.0;
    }
    /* pass B */
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nbB; ++b) {
      double lane[256];
      for (int t = 0; t < 256; ++t) lane[t] = 0.0;
      for (int j = 0; j < 32; ++j) {
        const int64_t g = b * 32 + j;
        double v = 0.0;
        if (g < G) {
          const int32_t s0 = o->offsets[g], s1 = o->offsets[g + 1];
          const double di = invD[o->gatherIds[s0]];
          for (int f = 0; f < Nf; ++f) {
            const int64_t fo = (int64_t)f * bp->offset;
            double rr = r[fo + o->gatherIds[s0]];
            if (k >= 0) {
              double ag = 0.0;
              for (int32_t q = s0; q < s1; ++q) ag = ag + Ap[fo + o->gatherIds[q]];
              rr = fma(-alpha, ag, rr);
              for (int32_t q = s0; q < s1; ++q) {
                const int64_t id = fo + o->gatherIds[q];
                x[id] = fma(alpha, p[id], x[id]);
                r[id] = rr;
              }
            }
            const double zz = di * rr;
            for (int32_t q = s0; q < s1; ++q) z[fo + o->gatherIds[q]] = zz;
            v = fma(rr, zz, v);
          }
        }
        lane[8 * j] = v;
      }
      pb[b] = nbo_tree256(lane);
    }
  }
  free(z); free(p); free(Ap); free(share); free(pa); free(pb);
  return 0;
}

/* V-cycle at level l: x = M_l r (x overwritten) */
void nbo_mg_vcycle(const nbo_mg* mg, int l, const double* r, double* x)
{
  if (mg->prec != 0 && l < mg->nlev - 1) {
    mg_vcycle_f(mg, l, r, x);
    return;
  }
  const nbo_bp* bp = &mg->bps[l];
  const int Nq = bp->N + 1, Np = Nq * Nq * Nq;
  const int64_t n = (int64_t)bp->E * Np;
  const int Nf = bp->Nf;
  const int64_t Nall = (int64_t)Nf * bp->offset;
  const double* invD = mg->invD[l];
  if (l == mg->nlev - 1) {  /* coarse: fixed-count Jacobi PCG from x = 0 (r is copied) */
    double* rr = (double*)malloc(sizeof(double) * Nall);
    memcpy(rr, r, sizeof(double) * Nall);
    if (bp->N != 1 || coarse_cg(bp, invD, rr, x, mg->coarse_iters) != 0) {
      memset(x, 0, sizeof(double) * Nall);
      nbo_pcg_stats st;
      pcg_core(bp, 1, 0, 1, NULL, invD, rr, x, 0.0, 0.0, mg->coarse_iters, 0, NULL, NULL, &st);
    }
    free(rr);
    return;
  }
  const nbo_bp* bc = &mg->bps[l + 1];
  const int Nqc = bc->N + 1, Npc = Nqc * Nqc * Nqc;
  const double* J = mg->J[l];
  const double* c1 = mg->c1 + (size_t)l * mg->degree;
  const double* c2 = mg->c2 + (size_t)l * mg->degree;
  const int k = mg->degree;
  double* res = (double*)calloc(Nall, sizeof(double));
  double* d = (double*)calloc(Nall, sizeof(double));
  double* Ad = (double*)calloc(Nall, sizeof(double));
  double* rc = (double*)calloc((size_t)Nf * bc->offset, sizeof(double));
  double* xc = (double*)calloc((size_t)Nf * bc->offset, sizeof(double));
  /* pre-smoothing from x = 0 */
  for (int f = 0; f < Nf; ++f)
    for (int64_t i = 0; i < n; ++i) {
      const int64_t id = i + (int64_t)f * bp->offset;
      const double ri = r[id];
      res[id] = ri;
      const double t = c2[0] * (invD[i] * ri);
      d[id] = t;
      x[id] = t;
    }
  for (int s = 1; s < k; ++s) {
    nbo_bp_apply(bp, d, Ad);
    for (int f = 0; f < Nf; ++f)
      for (int64_t i = 0; i < n; ++i) {
        const int64_t id = i + (int64_t)f * bp->offset;
        const double ri = res[id] - Ad[id];
        res[id] = ri;
        const double t = c2[s] * (invD[i] * ri);
        d[id] = fma(c1[s], d[id], t);
        x[id] = x[id] + d[id];
      }
  }
  nbo_bp_apply(bp, d, Ad);
  /* restriction: res -= Ad; rc = J^T (w res) per element and field, x then y then z */
#pragma omp parallel
  {
    double* u = (double*)malloc(sizeof(double) * Np);
    double* t1 = (double*)malloc(sizeof(double) * Np);
    double* t2 = (double*)malloc(sizeof(double) * Np);
#pragma omp for schedule(static) collapse(2)
    for (int f = 0; f < Nf; ++f)
      for (int32_t e = 0; e < bp->E; ++e) {
        const int64_t fo = (int64_t)f * bp->offset, fc = (int64_t)f * bc->offset;
        for (int q = 0; q < Np; ++q) {
          const int64_t id = (int64_t)e * Np + q;
          const double rq = res[fo + id] - Ad[fo + id];
          res[fo + id] = rq;
          u[q] = bp->w[id] * rq;
        }
        for (int q = 0; q < Nqc * Nq * Nq; ++q) {
          const int a = q % Nqc, jk = q / Nqc;
          double acc = 0.0;
          for (int i = 0; i < Nq; ++i) acc = fma(J[i * Nqc + a], u[i + Nq * jk], acc);
          t1[q] = acc;
        }
        for (int q = 0; q < Nqc * Nqc * Nq; ++q) {
          const int a = q % Nqc, b = (q / Nqc) % Nqc, kk = q / (Nqc * Nqc);
          double acc = 0.0;
          for (int j = 0; j < Nq; ++j) acc = fma(J[j * Nqc + b], t1[a + Nqc * (j + Nq * kk)], acc);
          t2[q] = acc;
        }
        for (int q = 0; q < Npc; ++q) {
          const int ab = q % (Nqc * Nqc), c = q / (Nqc * Nqc);
          double acc = 0.0;
          for (int kk = 0; kk < Nq; ++kk) acc = fma(J[kk * Nqc + c], t2[ab + Nqc * Nqc * kk], acc);
          rc[fc + (int64_t)e * Npc + q] = acc;
        }
      }
    free(u); free(t1); free(t2);
  }
  nbo_gs_many(bc->ogs, rc, Nf, bc->offset, NBO_DOUBLE, NBO_ADD);
  for (int f = 0; f < Nf; ++f)
    for (int32_t m = 0; m < bc->Nmasked; ++m) rc[bc->maskIds[m] + (int64_t)f * bc->offset] = 0.0;
  nbo_mg_vcycle(mg, l + 1, rc, xc);
  /* prolongation: d = J xc; x += d */
#pragma omp parallel
  {
    double* t1 = (double*)malloc(sizeof(double) * Np);
    double* t2 = (double*)malloc(sizeof(double) * Np);
#pragma omp for schedule(static) collapse(2)
    for (int f = 0; f < Nf; ++f)
      for (int32_t e = 0; e < bp->E; ++e) {
        const int64_t fo = (int64_t)f * bp->offset, fc = (int64_t)f * bc->offset;
        const double* uc = xc + fc + (int64_t)e * Npc;
        for (int q = 0; q < Nq * Nqc * Nqc; ++q) {
          const int i = q % Nq, bc_ = q / Nq;
          double acc = 0.0;
          for (int a = 0; a < Nqc; ++a) acc = fma(J[i * Nqc + a], uc[a + Nqc * bc_], acc);
          t1[q] = acc;
        }
        for (int q = 0; q < Nq * Nq * Nqc; ++q) {
          const int i = q % Nq, j = (q / Nq) % Nq, c = q / (Nq * Nq);
          double acc = 0.0;
          for (int b = 0; b < Nqc; ++b) acc = fma(J[j * Nqc + b], t1[i + Nq * (b + Nqc * c)], acc);
          t2[q] = acc;
        }
        for (int q = 0; q < Np; ++q) {
          const int ij = q % (Nq * Nq), kk = q / (Nq * Nq);
          double acc = 0.0;
          for (int c = 0; c < Nqc; ++c) acc = fma(J[kk * Nqc + c], t2[ij + Nq * Nq * c], acc);
          const int64_t id = fo + (int64_t)e * Np + q;
          d[id] = acc;
          x[id] = x[id] + acc;
        }
      }
    free(t1); free(t2);
  }
  nbo_bp_apply(bp, d, Ad);
  /* post-smoothing */
  for (int s = 0; s < k; ++s) {
    if (s > 0) nbo_bp_apply(bp, d, Ad);
    for (int f = 0; f < Nf; ++f)
      for (int64_t i = 0; i < n; ++i) {
        const int64_t id = i + (int64_t)f * bp->offset;
        const double ri = res[id] - Ad[id];
        res[id] = ri;
        const double t = c2[s] * (invD[i] * ri);
        const double di = s == 0 ? t : fma(c1[s], d[id], t);
        d[id] = di;
        x[id] = x[id] + di;
      }
  }
  if (l == 0 && mg->W) {  /* learned step: rf = res - A d */
    nbo_bp_apply(bp, d, Ad);
    for (int64_t i = 0; i < n; ++i) Ad[i] = res[i] - Ad[i];
    mg_learned_step(mg, Ad, x);
  }
  free(res); free(d); free(Ad); free(rc); free(xc);
}

/* BPPCG with the V-cycle as preconditioner (library: nbx_bp_pcg, precon MG):
 * flexible beta, the stopping rule of BPPCG */
int nbo_bp_pcg_mg(const nbo_mg* mg, double* r, double* x, double tol, int maxit, double* hist_rdotr,
                  double* hist_alpha, nbo_pcg_stats* st)
{
  const learned_ctx ln = {NULL, NULL, 0, 0.0, mg};
  return pcg_core_ext(&mg->bps[0], 1, 1, 1, NULL, mg->invD[0], &ln, r, x, tol, tol * tol, maxit, 1, hist_rdotr,
                      hist_alpha, st);
}

/* ---- FP32 cycle (library: vcycle<float> in nbx_mg.hip) -------------------
 * Every level but the last runs in float: Jacobi diagonal, interpolation,
 * Chebyshev coefficients rounded to float, FP32 operator (nbo_axhelm_f32 with
 * FP32 or bf16-valued geometry), float gather-scatter.  Level 0 reads the
 * double residual and accumulates the double correction; the transfers to
 * and from the last level (FP64 PCG) convert. */
static void mg_apply_f(const nbo_mg* mg, int l, const float* q, float* Aq)
{
  const nbo_bp* bp = &mg->bps[l];
  if (mg->prec == 3 && bp->N == 7)  /* matrix-core contractions at N = 7 (one field) */
    nbo_axhelm_bf16mma(bp->bk, bp->E, bp->offset, mg->ggeo32[l], mg->D32[l], mg->lam32[l], q, Aq);
  else
    nbo_axhelm_f32(bp->N, bp->Nf, bp->bk, bp->E, bp->offset, mg->ggeo32[l], mg->D32[l], mg->lam32[l], q, Aq);
  nbo_gs_many(bp->ogs, Aq, bp->Nf, bp->offset, NBO_FLOAT, NBO_ADD);
  for (int f = 0; f < bp->Nf; ++f)
    for (int32_t m = 0; m < bp->Nmasked; ++m) Aq[bp->maskIds[m] + (int64_t)f * bp->offset] = 0.f;
}

static void mg_vcycle_f(const nbo_mg* mg, int l, const void* r, void* x)
{
  const nbo_bp* bp = &mg->bps[l];
  const nbo_bp* bc = &mg->bps[l + 1];
  const int Nq = bp->N + 1, Np = Nq * Nq * Nq, Nqc = bc->N + 1, Npc = Nqc * Nqc * Nqc;
  const int64_t n = (int64_t)bp->E * Np;
  const int Nf = bp->Nf;
  const int64_t Nall = (int64_t)Nf * bp->offset, Ncall = (int64_t)Nf * bc->offset;
  const int rdbl = l == 0, cnext = l + 2 == mg->nlev;
  const float* invD = mg->invD32[l];
  const float* J = mg->J32[l];
  const int k = mg->degree;
  const double* c1 = mg->c1 + (size_t)l * k;
  const double* c2 = mg->c2 + (size_t)l * k;
  const double* rD = (const double*)r;
  const float* rF = (const float*)r;
  double* xD = (double*)x;
  float* xF = (float*)x;
  float* res = (float*)calloc(Nall, sizeof(float));
  float* d = (float*)calloc(Nall, sizeof(float));
  float* Ad = (float*)calloc(Nall, sizeof(float));
  const size_t ces = cnext ? sizeof(double) : sizeof(float);
  void* rc = calloc(Ncall, ces);
  void* xc = calloc(Ncall, ces);
#define XADD(i, v) do { if (rdbl) xD[i] = xD[i] + (double)(v); else xF[i] = xF[i] + (v); } while (0)
#define FIELDS for (int f = 0; f < Nf; ++f) for (int64_t i = 0; i < n; ++i)
  {
    const float c2f = (float)c2[0];
    FIELDS {
      const int64_t id = i + (int64_t)f * bp->offset;
      const float ri = rdbl ? (float)rD[id] : rF[id];
      res[id] = ri;
      const float t = c2f * (invD[i] * ri);
      d[id] = t;
      if (rdbl) xD[id] = (double)t; else xF[id] = t;
    }
  }
  for (int s = 1; s < k; ++s) {
    mg_apply_f(mg, l, d, Ad);
    const float c1f = (float)c1[s], c2f = (float)c2[s];
    FIELDS {
      const int64_t id = i + (int64_t)f * bp->offset;
      const float ri = res[id] - Ad[id];
      res[id] = ri;
      const float t = c2f * (invD[i] * ri);
      const float di = fmaf(c1f, d[id], t);
      d[id] = di;
      XADD(id, di);
    }
  }
  mg_apply_f(mg, l, d, Ad);
#pragma omp parallel
  {
    float* u = (float*)malloc(sizeof(float) * Np);
    float* t1 = (float*)malloc(sizeof(float) * Np);
    float* t2 = (float*)malloc(sizeof(float) * Np);
#pragma omp for schedule(static) collapse(2)
    for (int f = 0; f < Nf; ++f)
      for (int32_t e = 0; e < bp->E; ++e) {
        const int64_t fo = (int64_t)f * bp->offset, fc = (int64_t)f * bc->offset;
        for (int q = 0; q < Np; ++q) {
          const int64_t id = (int64_t)e * Np + q;
          const float rq = res[fo + id] - Ad[fo + id];
          res[fo + id] = rq;
          u[q] = (float)bp->w[id] * rq;
        }
        for (int q = 0; q < Nqc * Nq * Nq; ++q) {
          const int a = q % Nqc, jk = q / Nqc;
          float acc = 0.f;
          for (int i = 0; i < Nq; ++i) acc = fmaf(J[i * Nqc + a], u[i + Nq * jk], acc);
          t1[q] = acc;
        }
        for (int q = 0; q < Nqc * Nqc * Nq; ++q) {
          const int a = q % Nqc, b = (q / Nqc) % Nqc, kk = q / (Nqc * Nqc);
          float acc = 0.f;
          for (int j = 0; j < Nq; ++j) acc = fmaf(J[j * Nqc + b], t1[a + Nqc * (j + Nq * kk)], acc);
          t2[q] = acc;
        }
        for (int q = 0; q < Npc; ++q) {
          const int ab = q % (Nqc * Nqc), c = q / (Nqc * Nqc);
          float acc = 0.f;
          for (int kk = 0; kk < Nq; ++kk) acc = fmaf(J[kk * Nqc + c], t2[ab + Nqc * Nqc * kk], acc);
          if (cnext) ((double*)rc)[fc + (int64_t)e * Npc + q] = (double)acc;
          else ((float*)rc)[fc + (int64_t)e * Npc + q] = acc;
        }
      }
    free(u); free(t1); free(t2);
  }
  if (cnext) {
    nbo_gs_many(bc->ogs, rc, Nf, bc->offset, NBO_DOUBLE, NBO_ADD);
    for (int f = 0; f < Nf; ++f)
      for (int32_t m = 0; m < bc->Nmasked; ++m) ((double*)rc)[bc->maskIds[m] + (int64_t)f * bc->offset] = 0.0;
    nbo_mg_vcycle(mg, l + 1, (const double*)rc, (double*)xc);
  } else {
    nbo_gs_many(bc->ogs, rc, Nf, bc->offset, NBO_FLOAT, NBO_ADD);
    for (int f = 0; f < Nf; ++f)
      for (int32_t m = 0; m < bc->Nmasked; ++m) ((float*)rc)[bc->maskIds[m] + (int64_t)f * bc->offset] = 0.f;
    mg_vcycle_f(mg, l + 1, rc, xc);
  }
#pragma omp parallel
  {
    float* u = (float*)malloc(sizeof(float) * Npc);
    float* t1 = (float*)malloc(sizeof(float) * Np);
    float* t2 = (float*)malloc(sizeof(float) * Np);
#pragma omp for schedule(static) collapse(2)
    for (int f = 0; f < Nf; ++f)
      for (int32_t e = 0; e < bp->E; ++e) {
        const int64_t fo = (int64_t)f * bp->offset, fc = (int64_t)f * bc->offset;
        for (int q = 0; q < Npc; ++q)
          u[q] = cnext ? (float)((double*)xc)[fc + (int64_t)e * Npc + q] : ((float*)xc)[fc + (int64_t)e * Npc + q];
        for (int q = 0; q < Nq * Nqc * Nqc; ++q) {
          const int i = q % Nq, bc_ = q / Nq;
          float acc = 0.f;
          for (int a = 0; a < Nqc; ++a) acc = fmaf(J[i * Nqc + a], u[a + Nqc * bc_], acc);
          t1[q] = acc;
        }
        for (int q = 0; q < Nq * Nq * Nqc; ++q) {
          const int i = q % Nq, j = (q / Nq) % Nq, c = q / (Nq * Nq);
          float acc = 0.f;
          for (int b = 0; b < Nqc; ++b) acc = fmaf(J[j * Nqc + b], t1[i + Nq * (b + Nqc * c)], acc);
          t2[q] = acc;
        }
        for (int q = 0; q < Np; ++q) {
          const int ij = q % (Nq * Nq), kk = q / (Nq * Nq);
          float acc = 0.f;
          for (int c = 0; c < Nqc; ++c) acc = fmaf(J[kk * Nqc + c], t2[ij + Nq * Nq * c], acc);
          const int64_t id = fo + (int64_t)e * Np + q;
          d[id] = acc;
          XADD(id, acc);
        }
      }
    free(u); free(t1); free(t2);
  }
  mg_apply_f(mg, l, d, Ad);
  for (int s = 0; s < k; ++s) {
    if (s > 0) mg_apply_f(mg, l, d, Ad);
    const float c1f = (float)c1[s], c2f = (float)c2[s];
    FIELDS {
      const int64_t id = i + (int64_t)f * bp->offset;
      const float ri = res[id] - Ad[id];
      res[id] = ri;
      const float t = c2f * (invD[i] * ri);
      const float di = s == 0 ? t : fmaf(c1f, d[id], t);
      d[id] = di;
      XADD(id, di);
    }
  }
  if (rdbl && mg->W) {  /* learned step on the fine level: rf = res - A d, as double */
    mg_apply_f(mg, l, d, Ad);
    double* rf = (double*)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i) rf[i] = (double)(res[i] - Ad[i]);
    mg_learned_step(mg, rf, xD);
    free(rf);
  }
#undef FIELDS
#undef XADD
  free(res); free(d); free(Ad); free(rc); free(xc);
}
