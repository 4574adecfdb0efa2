/*
 * nbo -- CPU oracle for the nekBench solver hot path (axhelm, ogs
 * gather-scatter, nekBone PCG).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark compare against.
 * Nothing in the product (nekbench_amd/, include/) links or calls it.
 *
 * Every function restates a reference routine; the reference file:line it
 * follows is given at each declaration (paths relative to the nekBench tree).
 *
 * Floating-point contract: the library is compiled with -ffp-contract=off,
 * so every a*b+c below is two roundings unless written as fma().  The FP64
 * GPU kernels use the identical operation order (fma where this file uses
 * fma), which makes bit-exact comparison possible.
 */
#ifndef NBO_H
#define NBO_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- counter-based PRNG (splitmix64) ------------------------------------ */
uint64_t nbo_splitmix64(uint64_t x);
/* uniform in [-1,1): value i of stream `stream` under `seed` */
double nbo_urand(uint64_t seed, uint64_t stream, uint64_t i);
void nbo_urand_fill(uint64_t seed, uint64_t stream, int64_t n, double* out);

/* ---- 1-D GLL basis (core/meshBasis.cpp:140-176, 775-790) ---------------- */
void nbo_gll(int N, double* z, double* w);                 /* z, w: N+1 */
void nbo_dmatrix(int N, const double* z, double* D);       /* D[i*Nq+j] = l_j'(z_i) */

/* ---- box mesh (core/meshSetup.cpp:2265-2451, 2156-2233, 215-360, 1638-1735) */
typedef struct {
  int N;                 /* polynomial degree */
  int NX, NY, NZ;        /* elements per direction */
  int periodic;          /* 1: periodic box (reference nekBone), 0: bounded box */
  double perturb;        /* interior vertex noise, fraction of local element width */
  double zaspect;        /* 0: uniform z; >0: geometric z grading, first layer = hx/zaspect */
  uint64_t seed;
  double x0, x1, y0, y1, z0, z1;  /* box extent */
} nbo_box;

int64_t nbo_box_nelements(const nbo_box* b);
/* element vertex coordinates EX/EY/EZ[e*8+v] in the reference vertex order and
 * EToV[e*8+v] (global vertex number).  Elements are numbered e = ex + NX*(ey + NY*ez). */
void nbo_box_vertices(const nbo_box* b, double* EX, double* EY, double* EZ, int64_t* EToV);
/* trilinear map of the GLL nodes (core/meshSetup.cpp:2156-2233) */
void nbo_physical_nodes(int N, int64_t E, const double* EX, const double* EY, const double* EZ,
                        const double* z, double* x, double* y, double* zc);
/* ggeo[e][7][Np] (core/meshSetup.cpp:266-360); returns min J */
double nbo_geometric_factors(int N, int64_t E, const double* x, const double* y, const double* zc,
                             const double* D, const double* w, double* ggeo);
/* global ids as meshParallelConnectNodes leaves them (core/meshSetup.cpp:1638-1735):
 * vertex nodes carry EToV+1, other nodes 1+Nnodes+min local id of their group.
 * mask_boundary != 0 zeroes the ids of nodes on the box boundary (Dirichlet). */
void nbo_global_ids(const nbo_box* b, int mask_boundary, int64_t* ids);

/* ---- ogs (3rdParty/ogs/src/ogsSetup.cpp:42-212, 325-352) ----------------- */
typedef struct {
  int32_t N;             /* local nodes */
  int32_t Nlocal;        /* nodes with id != 0 */
  int32_t NlocalGather;  /* number of groups */
  int32_t* offsets;      /* [NlocalGather+1] */
  int32_t* gatherIds;    /* [Nlocal] */
  double* invDegree;     /* [N] */
} nbo_ogs;
int nbo_ogs_setup(int32_t N, const int64_t* ids, nbo_ogs* out);
void nbo_ogs_free(nbo_ogs* o);

enum { NBO_DOUBLE = 0, NBO_FLOAT = 1, NBO_INT = 2, NBO_LONG = 3 };
enum { NBO_ADD = 0, NBO_MUL = 1, NBO_MIN = 2, NBO_MAX = 3 };
/* gatherScatterMany_<type><op> (3rdParty/ogs/okl/gatherScatterMany.okl:58-86) */
int nbo_gs_many(const nbo_ogs* o, void* q, int k, int32_t stride, int type, int op);

/* ---- axhelm (axhelm/kernel/CPU/axhelm.c:27-413; axhelm/axhelmReference.cpp) */
/* Nf in {1,3}; bk != 0 ignores lambda.  elementList == NULL: all elements
 * (axhelm_*), else the Partial variants over elementList[0..Nelements).
 * 1 <= N <= 15; q and Aq must not overlap (the reference's @restrict). */
void nbo_axhelm(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const int32_t* elementList,
                const double* ggeo, const double* D, const double* lambda,
                const double* q, double* Aq);
/* matrix-core restatements, N = 7, one field: bf16 operands (ggeo and lambda
 * hold bf16 values) or fp16 operands (FP32 ggeo and lambda) */
void nbo_axhelm_bf16mma(int bk, int32_t E, int32_t offset, const float* ggeo, const float* D, const float* lambda,
                        const float* q, float* Aq);
void nbo_axhelm_f16mma(int bk, int32_t E, int32_t offset, const float* ggeo, const float* D, const float* lambda,
                       const float* q, float* Aq);
void nbo_axhelm_mma_n(int f16, int N, int Nf, int bk, int32_t E, int32_t offset, const float* ggeo, const float* D,
                      const float* lambda, const float* q, float* Aq);
/* IEEE binary16 rounding (nearest even, subnormals kept) of each value */
void nbo_f16_round_array(int64_t n, const float* x, float* y);
/* bf16 rounding of a float (the value of its upper 16 bits after round to
 * nearest, ties to even; past the largest finite bf16 -> infinity; a NaN stays
 * a NaN, quieted).  Every bf16 operand of the oracle goes through it. */
float nbo_bf16r(float x);
void nbo_bf16_round_array(int64_t n, const float* x, float* y);
/* single-precision restatement (same loop nest, float arithmetic) */
void nbo_axhelm_f32(int N, int Nf, int bk, int32_t Nelements, int32_t offset,
                    const float* ggeo, const float* D, const float* lambda,
                    const float* q, float* Aq);
/* axhelmReference (axhelm/axhelmReference.cpp:19-101): lambda*GWJ*q + D^T G D q */
void nbo_axhelm_reference(int Nq, int numElements, double lambda, const double* ggeo,
                          const double* D, const double* solIn, double* solOut);

/* ---- Jacobi (nekBone/kernel/updateJacobi.okl:27-108) --------------------- */
void nbo_jacobi_local(int N, int32_t Nelements, int32_t offset, const int32_t* mapB,
                      const double* ggeo, const double* D, const double* lambda, double* diag);

/* ---- reductions with the library's fixed tree ---------------------------- */
/* tile = 256 lanes x 8 values; per lane sequential fma, then the 256-lane tree
 * (64-lane halving per wave, then (S0+S2)+(S1+S3)); stage 2 strides partials
 * over 256 lanes, adds sequentially and applies the same tree. */
double nbo_tree_wdot(int64_t n, const double* w, const double* x, const double* y);
double nbo_tree_dot(int64_t n, const double* x, const double* y);
double nbo_tree_sum_partials(int64_t nb, const double* partials);
double nbo_tree256(const double* lanes);  /* the 256-lane tree alone */
int nbo_max_threads(void);                 /* OpenMP threads the oracle uses */

/* ---- BP operator and PCG (nekBone/solve.cpp:44-293) --------------------- */
typedef struct {
  int N, Nf, bk;
  int32_t E, offset;          /* offset = field stride = E*Np */
  const double *ggeo, *D, *lambda;
  const nbo_ogs* ogs;         /* gather-scatter of the (masked) ids */
  const double* w;            /* invDegree */
  int32_t Nmasked;
  const int32_t* maskIds;
} nbo_bp;

typedef struct {
  int iterations, converged, breakdown, iter;
  double rdotr0, rdotr, TOL;
  int fallback_iter;
} nbo_pcg_stats;

/* Aq = mask(gs(axhelm(q))) per field (BPAxOperator, solve.cpp:181-235 without the dot) */
void nbo_bp_apply(const nbo_bp* bp, const double* q, double* Aq);
/* <x,y>_w over the fields with the library's tree (weightedMultipleInnerProduct2) */
double nbo_bp_wdot(const nbo_bp* bp, const double* x, const double* y);
/* BPPCG (solve.cpp:44-113) in the library's order: returns iterations (iter-1).
 * jacobi: z = invDiagA r with invDiagA = 1/gs(local diag) (solve.cpp:146-165);
 * tol <= 0 runs maxit iterations (FIXED ITERATION COUNT).  hist_rdotr[it],
 * hist_alpha[it] (may be NULL, length >= maxit+2) record each iteration.
 * fused: pAp from the local operator output (the library's fused schedule). */
int nbo_bp_pcg(const nbo_bp* bp, int jacobi, int flexible, int fused, double* r, double* x, double tol, int maxit,
               double* hist_rdotr, double* hist_alpha, nbo_pcg_stats* st);

/* Mixed precision with iterative refinement (library: nbx_bp_pcg_mixed):
 * prec 0 = FP64, 1 = FP32 operator, 2 = bf16 ggeo/lambda with FP32 arithmetic.
 * st->iter holds the number of refinement cycles. */
int nbo_bp_pcg_mixed(const nbo_bp* bp, int jacobi, int flexible, int prec, double* r, double* x, double tol,
                     int maxit, double inner_tol, int max_cycles, nbo_pcg_stats* st);

void nbo_bp_zero_mean(const nbo_bp* bp, double* q);

/* ---- p-multigrid preconditioner (nbo_pcg.c) ------------------------------ */
typedef struct nbo_mg {
  int nlev;
  const nbo_bp* bps;           /* levels, finest first */
  const double* const* invD;   /* assembled inverse Jacobi diagonal per level */
  const double* const* J;      /* level l < nlev-1: [N_l+1][N_{l+1}+1] interpolation */
  int degree;                  /* Chebyshev degree */
  const double* c1;            /* [nlev][degree] Chebyshev coefficients */
  const double* c2;
  int coarse_iters;            /* Jacobi-PCG iterations on the last level */
  int prec;                    /* 0: FP64 cycle; 1 FP32, 2 bf16 geometry, 3 bf16 geometry with bf16
                                  matrix-core contractions at N = 7: FP32 cycle but the last level */
  const float* const* ggeo32;  /* per level: ggeo as FP32 (or bf16-rounded) values */
  const float* const* lam32;   /* per level: lambda (2 offset) likewise */
  const float* const* D32;
  const float* const* invD32;
  const float* const* J32;
  /* learned step (library: nbx_bp_mg_set_learned): when W != NULL the fine
   * level ends the cycle with x += mask(gs(net(res - A d))), the element
   * network of nbo_net_apply with weights W and features feat */
  const float* W;
  const float* feat;
} nbo_mg;
void nbo_mg_interp(int Nf, int Nc, double* J);
void nbo_mg_cheb(double lmax, double lo, double hi, int degree, double* c1, double* c2);
double nbo_mg_power(const nbo_bp* bp, const double* invD, uint64_t seed, uint64_t stream, int iters);
void nbo_mg_vcycle(const nbo_mg* mg, int l, const double* r, double* x);
int nbo_bp_pcg_mg(const nbo_mg* mg, double* r, double* x, double tol, int maxit, double* hist_rdotr,
                  double* hist_alpha, nbo_pcg_stats* st);

/* ---- learned element preconditioner (nbo_net.c) -------------------------- */
/* zl = w invDiag (r + s f(r)) per element (w NULL: weight 1) */
void nbo_net_apply(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                   const double* w, double* zl);
void nbo_net_features(int32_t E, int Np, const double* ggeo, float* feat);
void nbo_net_apply_mma(int32_t E, const float* W, const float* feat, const double* r, const double* invDiag,
                       const double* w, double* zl);
double nbo_sumsq_inv_f(int64_t n, const float* a);
void nbo_net_gradient(int32_t nsamples, int32_t chunk, const int32_t* samples, int64_t pairStride,
                      const float* R, const float* Ev, const double* invNorm, const float* W, const float* feat,
                      const double* invDiag, const double* Gz, float* grad, double* loss);
int nbo_bp_pcg_learned(const nbo_bp* bp, const float* W, const float* feat, int window, double ratio, double* r,
                       double* x, double tol, int maxit, double* hist_rdotr, double* hist_alpha, nbo_pcg_stats* st);
/* nbo_bp_pcg_mixed with the learned preconditioner in every inner solve */
int nbo_bp_pcg_mixed_learned(const nbo_bp* bp, const float* W, const float* feat, int window, double ratio, int prec,
                             double* r, double* x, double tol, int maxit, double inner_tol, int max_cycles,
                             nbo_pcg_stats* st);

#ifdef __cplusplus
}
#endif
#endif
