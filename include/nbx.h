/*
 * nbx.h -- C ABI of the MI355X (gfx950) nekBench solver hot path.
 *
 * Three drop-in boundaries, each replacing one reference interface:
 *   1. axhelm kernels      <- loadAxKernel + axKernel(Nelements, offset, ggeo, D, lambda, q, Aq)
 *   2. ogs gather-scatter  <- ogsSetup / ogsGatherScatterMany / oogs::start+finish
 *   3. BP driver           <- setup / BPPCG / BPAxOperator of nekBone
 *
 * Conventions (shared by every entry point):
 *   - plain C types only; all array arguments are device pointers unless the
 *     comment says "host";
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream);
 *     work is enqueued asynchronously on it, as the reference enqueues on the
 *     OCCA device's current stream;
 *   - the caller owns every buffer it passes; handles own what they allocate;
 *   - return value 0 = success, < 0 = error (see NBX_ERR_*); nbx_last_error()
 *     gives the message.  No entry point calls exit();
 *   - dlong is int32 (reference makefile:40, dlong=int), hlong is int64
 *     (makefile:45); ggeo and element offsets are indexed in 64 bits inside.
 *   - layout: element-major, node index i + j*Nq + k*Nq^2
 *     (axhelm/kernel/CPU/axhelm.c:52); ggeo[e][7][Np] with factor order
 *     GWJ, G00, G01, G02, G11, G12, G22 (core/params.h:55-62); field f of a
 *     multi-field vector at f*offset.
 */
#ifndef NBX_H
#define NBX_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t nbx_dlong;
typedef int64_t nbx_hlong;

#define NBX_SUCCESS 0
#define NBX_ERR_INVALID (-1)     /* bad argument */
#define NBX_ERR_UNSUPPORTED (-2) /* combination not built (e.g. N > 15) */
#define NBX_ERR_HIP (-3)         /* HIP runtime error */
#define NBX_ERR_ALLOC (-4)       /* allocation failed */
#define NBX_ERR_DIVERGED (-5)    /* PCG breakdown (non-finite or non-positive curvature) */

/* NBX_BF16_MFMA: bf16 geometry and bf16 MFMA tensor contractions with FP32
 * accumulation, FP32 vectors (nbx_axhelm: N = 7 or 9, one or three fields, no
 * elementList).  NBX_F16: FP32 geometry and fp16 MFMA tensor contractions with
 * FP32 accumulation, FP32 vectors (nbx_axhelm only, same limits). */
typedef enum { NBX_F64 = 0, NBX_F32 = 1, NBX_BF16 = 2, NBX_F16 = 3, NBX_BF16_MFMA = 4 } nbx_prec;
/* ogs types / ops (3rdParty/ogs/ogs.hpp:128-140: "double", "float", "int",
 * "long long int"; "add", "mul", "min", "max") */
typedef enum { NBX_DOUBLE = 0, NBX_FLOAT = 1, NBX_INT = 2, NBX_LONG = 3 } nbx_type;
typedef enum { NBX_ADD = 0, NBX_MUL = 1, NBX_MIN = 2, NBX_MAX = 3 } nbx_op;

const char* nbx_last_error(void);
int nbx_version(void);
int nbx_device_synchronize(void);

/* ======================= 1. axhelm ======================================
 * Replaces the kernels loaded by loadAxKernel (axhelm/kernelHelper.cpp:3-58),
 * called as kernel(Nelements, offset, ggeo, D, lambda, q, Aq)
 * (axhelm/main.cpp:132, nekBone/solve.cpp:218) and, for the Partial forms,
 * kernel(Nelements, offset, elementList, ggeo, D, lambda, q, Aq)
 * (axhelm/kernel/VOLTA/axhelm.okl:158-165, nekBone/solve.cpp:192-212).
 *
 *   Aq = D^T (lambda0 . G D q) + lambda1 . GWJ . q      (bk = 0, axhelm_v0)
 *   Aq = D^T (G D q)                                    (bk = 1, axhelm_bk_v0)
 * for Nfields = 1 or 3 (_n3: the same ggeo and lambda for every field).
 * lambda[0..offset) is lambda0, lambda[offset..2 offset) lambda1 (per node),
 * ignored when bk = 1 (may be NULL then).  q and Aq must not overlap; Aq is
 * overwritten.  elementList == NULL processes elements 0..Nelements-1.
 * prec selects the arithmetic and the element type of ggeo, D, lambda, q, Aq:
 * NBX_F64 (double), NBX_F32 (float), NBX_BF16 (float D, q, Aq and FP32
 * arithmetic; ggeo and lambda as bf16, i.e. the upper 16 bits of the float,
 * uint16_t; one field, no elementList), NBX_BF16_MFMA (as NBX_BF16, with q,
 * D and the G-scaled gradients rounded to bf16 for the matrix-core
 * contractions, FP32 accumulation; N = 7 or 9, one or three fields, no
 * elementList), NBX_F16 (float ggeo, D, lambda, q, Aq; q, D and the G-scaled
 * gradients rounded to fp16, nearest even, for the matrix-core contractions,
 * FP32 accumulation; N = 7 or 9, one or three fields, no elementList; fp16's
 * range suits q of order one).  N in [1, 15]. */
int nbx_axhelm(int N, int Nfields, int bk, nbx_prec prec,
               nbx_dlong Nelements, nbx_dlong offset, const nbx_dlong* elementList,
               const void* ggeo, const void* D, const void* lambda,
               const void* q, void* Aq, void* stream);

/* Same, with the kernel chosen by the reference's name grammar
 * "axhelm[Partial][_bk][_n3]_v<k>" (nekBone/setup.cpp:313-320,
 * axhelm/main.cpp:105-109), FP64.  Partial names require elementList. */
int nbx_axhelm_by_name(const char* kernelName, int N,
                       nbx_dlong Nelements, nbx_dlong offset, const nbx_dlong* elementList,
                       const double* ggeo, const double* D, const double* lambda,
                       const double* q, double* Aq, void* stream);

/* Local Jacobi diagonal of lambda0*A + lambda1*B (nekBone/kernel/updateJacobi.okl:27-108);
 * mapB (may be NULL) marks nodes whose diagonal is 1. */
int nbx_update_jacobi_local(int N, nbx_dlong Nelements, nbx_dlong offset, const int32_t* mapB,
                            const double* ggeo, const double* D, const double* lambda,
                            double* diagA, void* stream);

/* ======================= 2. ogs gather-scatter ==========================
 * Replaces ogsSetup (3rdParty/ogs/src/ogsSetup.cpp:68-355) for one rank and
 * ogsGatherScatterMany / oogs::start+finish
 * (3rdParty/ogs/src/ogsGatherScatterMany.cpp:54-183, oogs.cpp:248-342).
 * ids (host, length N): nodes sorted by (|id|, local index) with id 0
 * excluded; a group is a run of equal signed id in that order (ogsSetup.cpp:163;
 * all-positive ids simply group by value).  Groups and member order are exactly
 * those of ogsSetup: groups in order of first appearance by local index,
 * members ascending.  The
 * gather-scatter combines the members of every group of two or more nodes
 * in that order, starting from the op's identity, and writes the result to
 * each member; singletons and excluded nodes are left untouched
 * (3rdParty/ogs/okl/gatherScatterMany.okl:58-86). */
typedef struct nbx_ogs nbx_ogs;
nbx_ogs* nbx_ogs_setup(nbx_dlong N, const nbx_hlong* ids, int verbose);
/* Same, with ids in device memory (for ids produced on the GPU, e.g. by
 * nbx_box_mesh_device).  The CSR is built on the GPU in both forms: stable
 * radix sort of (|id|, local index), then scans. */
nbx_ogs* nbx_ogs_setup_device(nbx_dlong N, const nbx_hlong* ids, int verbose);
int nbx_ogs_gather_scatter_many(nbx_ogs* ogs, void* v, int k, nbx_dlong stride,
                                nbx_type type, nbx_op op, void* stream);
int nbx_ogs_gather_scatter(nbx_ogs* ogs, void* v, nbx_type type, nbx_op op, void* stream);
/* 1/multiplicity per local node, 1 for excluded nodes (ogsSetup.cpp:325-352); device */
const double* nbx_ogs_inv_degree(const nbx_ogs* ogs);
/* sizes: N, Nlocal (ids != 0), NlocalGather (groups), groups with >= 2 members,
 * nodes in groups with >= 2 members (host outputs, any may be NULL) */
int nbx_ogs_sizes(const nbx_ogs* ogs, nbx_dlong* N, nbx_dlong* Nlocal, nbx_dlong* NlocalGather,
                  nbx_dlong* NsharedGather, nbx_dlong* Nshared);
/* host copies of localGatherOffsets[NlocalGather+1] and localGatherIds[Nlocal] */
int nbx_ogs_copy_csr(const nbx_ogs* ogs, nbx_dlong* offsets, nbx_dlong* ids);
void nbx_ogs_free(nbx_ogs* ogs);

/* Dirichlet mask: q[maskIds[n]] = 0 (nekBone/kernel/utils.okl:127-134) */
int nbx_mask(nbx_dlong Nmasked, const nbx_dlong* maskIds, double* q, void* stream);

/* ======================= mesh inputs (host) ==============================
 * Produce the hot path's inputs as the reference's mesh setup does
 * (core/meshSetup.cpp:2265-2451).  All arrays here are HOST memory. */
typedef struct {
  int N;                       /* polynomial degree */
  int NX, NY, NZ;              /* elements per direction */
  int periodic;                /* 1: periodic box as the reference's nekBone */
  double perturb;              /* interior vertex noise, fraction of element width */
  double zaspect;              /* 0: uniform z; > 0: geometric z grading, first layer hx/zaspect */
  uint64_t seed;
  double x0, x1, y0, y1, z0, z1;
} nbx_box;

/* GLL nodes and weights (core/meshBasis.cpp:140-176) and D[i*Nq+j] = l_j'(z_i)
 * (core/meshBasis.cpp:775-790) */
int nbx_gll(int N, double* z, double* w, double* D);
/* uniform [-1,1) stream, counter-based splitmix64 */
int nbx_urand_fill(uint64_t seed, uint64_t stream, int64_t n, double* out);
/* Element vertices, nodes (x, y, z may be NULL), ggeo and global ids
 * (ids may be NULL; mask_boundary zeroes boundary ids for Dirichlet).
 * Returns the minimum Jacobian through minJ (may be NULL). */
int nbx_box_mesh(const nbx_box* box, int mask_boundary, double* x, double* y, double* z,
                 double* ggeo, nbx_hlong* globalIds, double* minJ);
/* nbx_box_mesh's ggeo and global ids computed on the GPU into device arrays
 * (either may be NULL), equal to the host ones bit for bit; synchronous on
 * `stream`.  (core/meshSetup.cpp:266-360, :1638-1735 on the device.) */
int nbx_box_mesh_device(const nbx_box* box, int mask_boundary, double* ggeo, nbx_hlong* globalIds, void* stream);

/* ======================= BLAS-1 of the PCG loop ==========================
 * nekBone/kernel/utils.okl.  Reductions return their value through a device
 * (result_dev) or host (result_host) pointer; results are deterministic: a
 * fixed tree independent of timing (tiles of 256 lanes x 8 values, lane
 * fma chain, per-wave halving tree, (S0+S2)+(S1+S3), partials combined the
 * same way by a one-workgroup second pass).  N = 0 gives 0 (the empty sum). */
int nbx_weighted_inner_product(nbx_dlong N, nbx_dlong offset, int Nfields, const double* w,
                               const double* x, const double* y, double* result_host, void* stream);
int nbx_scaled_add(nbx_dlong N, double alpha, const double* x, double beta, double* y, void* stream);
/* No counterpart in the reference.  nbx_weighted_inner_product for callers
 * that cannot allocate device memory (work under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the same value, bit for
 * bit, written to result_dev (device memory, one double) on `stream` without
 * a host wait, its tile partials in a fixed device array of the library
 * (512 KiB per GPU, kept until the process ends) instead of a library
 * allocation.  N <= NBX_WDOT_NOALLOC_MAX_N (larger N: NBX_ERR_INVALID).  All
 * calls on one GPU share that array: calls that overlap in time on one GPU
 * (concurrent streams, host threads) race and must be serialised by the
 * caller. */
#define NBX_WDOT_NOALLOC_MAX_N 134217728
int nbx_weighted_inner_product_noalloc(nbx_dlong N, nbx_dlong offset, int Nfields, const double* w,
                                       const double* x, const double* y, double* result_dev, void* stream);

/* Bandwidth probes of the bw benchmark (bw/main.cpp:59-140, bw/kernel/bwKernels.okl)
 * plus the copy and read streams: the measured HBM and LDS ceilings.  Device
 * arrays, 16-byte aligned (hipMalloc'd); asynchronous on `stream`.
 *   nbx_triad:     c = a + alpha b                (bwKernels.okl:1-11 "triad")
 *   nbx_copy:      c = a
 *   nbx_read_sum:  work[0] = sum of a[0..N) in a fixed order; work is a device
 *                  buffer of NBX_READ_SUM_WORK doubles (block partials after
 *                  work[0]); calls sharing a buffer must not overlap
 *   nbx_lds_probe: per block of 128 values, 128 rounds of s[(i+j)%128] += s[0]
 *                  through LDS (bwKernels.okl:13-42 "smem"); a[b*128+i] becomes
 *                  a[b*128+i] + 128 sequential additions of a[b*128] */
int nbx_triad(nbx_dlong N, double alpha, const double* a, const double* b, double* c, void* stream);
int nbx_copy(nbx_dlong N, const double* a, double* c, void* stream);
#define NBX_READ_SUM_WORK 2049
int nbx_read_sum(nbx_dlong N, const double* a, double* work, void* stream);
int nbx_lds_probe(nbx_dlong Nblocks, double* a, void* stream);

/* ======================= 3. BP driver ====================================
 * Replaces nekBone's setup()/solveSetup() (nekBone/setup.cpp:36-367),
 * BPPCG (nekBone/solve.cpp:44-113) and BPAxOperator (solve.cpp:181-235) on
 * one GPU.  The handle owns the mesh (D, ggeo, gather-scatter), the Jacobi
 * diagonal and the PCG workspace; r and x are the caller's. */
typedef struct nbx_bp nbx_bp;

/* COPY and JACOBI as nekBone's PRECONDITIONER option (solve.cpp:167-179);
 * JACOBI also for Nfields = 3 (one diagonal for every field; nekBone stops
 * there, setup.cpp:158-162).  LEARNED: the element network below; MG: the
 * p-multigrid V-cycle below.  LEARNED and MG force the flexible beta. */
typedef enum { NBX_PRECON_COPY = 0, NBX_PRECON_JACOBI = 1, NBX_PRECON_LEARNED = 2, NBX_PRECON_MG = 3 } nbx_precon;

typedef struct {
  nbx_box box;
  int Nfields;          /* 1 or 3 */
  int bpmode;           /* 1: BK kernels (BP5/BP6), lambda1 forced to 0 (setup.cpp:43-53) */
  double lambda1;       /* Helmholtz shift (non-BK) */
  int dirichlet;        /* 1: zero boundary nodes (ids 0 + mask) */
  nbx_precon precon;
  int flexible;         /* "PCG+FLEXIBLE": beta = -alpha (z.Ap)/rdotz_old (solve.cpp:92-95) */
  int verbose;
  int rhs_kind;         /* 0: U[-1,1) per local node (seed stream 3+field), masked, assembled;
                           1: (3 pi^2 + lambda1) JW cos(pi x) cos(pi y) cos(pi z), assembled
                              (nekBone/setup.cpp:77-99) */
  int fused;            /* 0: the reference's kernel sequence (pAp = <p, gs(A p)>_w);
                           1: p-update and pAp = sum_local p (A_L p) fused into the operator
                              pass, gather-scatter and mask in one launch (same algorithm,
                              pAp rounded differently) */
} nbx_bp_options;

typedef struct {
  int iterations;       /* BPPCG's return value (iter - 1) */
  int converged;
  double rdotr0;        /* weighted |r|^2 at iteration 1 */
  double rdotr;         /* weighted |r|^2 at exit */
  double elapsed_s;     /* wall time of the solve on the device */
  int cycles;           /* refinement cycles (1 for nbx_bp_pcg) */
  int fallback_iteration; /* learned preconditioner: iteration of the switch to Jacobi (0 = none) */
} nbx_pcg_stats;

nbx_bp* nbx_bp_setup(const nbx_bp_options* opt, void* stream);
void nbx_bp_free(nbx_bp* bp);
/* sizes and device arrays owned by the handle:
 * name in {"ggeo","D","lambda","invDegree","invDiagA","rhs","maskIds"} */
int nbx_bp_sizes(const nbx_bp* bp, nbx_dlong* Nelements, nbx_dlong* Np, nbx_dlong* fieldOffset,
                 nbx_dlong* Nmasked);
const void* nbx_bp_device_array(const nbx_bp* bp, const char* name);
nbx_ogs* nbx_bp_ogs(nbx_bp* bp);
/* BPAxOperator: Aq = gs(axhelm(q)) (+ mask); returns pAp = <q, Aq>_w through pAp_host */
int nbx_bp_ax_operator(nbx_bp* bp, const double* lambda, const double* q, double* Aq, double* pAp_host);
/* BPZeroMean (nekBone/solve.cpp:115-134): q -= unweighted mean of the
 * E-vector over its Nf fields, on the handle's stream.  nbx_bp_pcg applies it
 * to r before and x after the solve when nbx_bp_all_neumann() is 1: lambda1
 * is 0 (setup.cpp:210-214) and no Dirichlet mask is set (the reference never
 * masks, so it projects every lambda1 = 0 problem). */
int nbx_bp_zero_mean(nbx_bp* bp, double* q);
/* Right-hand side built as the setup builds "rhs" (U[-1,1) per local node
 * from splitmix64(seed, stream 3 + field), masked, assembled; nekBone/setup.cpp
 * :77-99) but with another seed; out is a device vector of Nfields*offset. */
int nbx_bp_make_rhs(nbx_bp* bp, uint64_t seed, double* out);
int nbx_bp_all_neumann(const nbx_bp* bp);

/* BPPCG: solves A x = r in place (r becomes the residual), tol relative,
 * stopping test rdotr <= max(tol^2 rdotr_1, tol^2) (solve.cpp:85-90). */
int nbx_bp_pcg(nbx_bp* bp, const double* lambda, double* r, double* x, double tol, int maxit,
               nbx_pcg_stats* stats);
/* Mixed-precision solve with iterative refinement: the true residual
 * r = b - A x is formed with the FP64 operator; each cycle solves A_lo d = r
 * by PCG from d = 0 to the relative tolerance inner_tol with the operator in
 * op_prec (NBX_F64, NBX_F32, or NBX_BF16 = bf16 ggeo with FP32 arithmetic)
 * while the vectors, dot products, gather-scatter sums, updates and the
 * stopping test stay FP64; then x += d.  Stops when |r|_w <= tol |r_0|_w
 * (true residual), after max_cycles cycles or maxit inner iterations in
 * total.  Returns the total inner iterations; stats->rdotr is the final
 * true residual.  The singular all-Neumann problem (nbx_bp_all_neumann) is
 * refused with NBX_ERR_UNSUPPORTED: its true residual need not reach tol. */
int nbx_bp_pcg_mixed(nbx_bp* bp, const double* lambda, double* r, double* x, double tol, int maxit,
                     nbx_prec op_prec, double inner_tol, int max_cycles, nbx_pcg_stats* stats);
/* ---- p-multigrid preconditioner (NBX_PRECON_MG) ---------------------------
 * No counterpart in the reference (nekBone offers COPY and JACOBI only).
 * Levels are the same mesh at orders orders[0] = N > orders[1] > ... >= 1;
 * one application is a V-cycle with Chebyshev-Jacobi smoothing of the given
 * degree on bounds [lo, hi] * lambda_max(D^-1 A) (power iteration at setup),
 * element-local GLL interpolation between levels, and coarse_iters Jacobi-PCG
 * iterations on the last level.  Used inside flexible PCG by nbx_bp_pcg and
 * nbx_bp_pcg_mixed when the handle was set up with precon = NBX_PRECON_MG.
 * prec: NBX_F64 (all FP64), NBX_F32, NBX_BF16 or NBX_BF16_MFMA (smoothing and
 * transfers of every level but the last in FP32, operator geometry in FP32 or
 * bf16, the N = 7 level's contractions on the matrix cores for MFMA; the
 * cycle reads the FP64 residual and returns an FP64 correction; the coarse
 * PCG is FP64).  One or three fields (one Jacobi diagonal and lambda_max for
 * all fields; bf16 geometry: one field), N <= 11.  Uses the handle's lambda.
 * Not for the singular all-Neumann problem (nbx_bp_all_neumann). */
int nbx_bp_mg_setup(nbx_bp* bp, int nlevels, const int* orders, int degree, int coarse_iters, double lo, double hi,
                    nbx_prec prec);
/* on (default): the cycle is captured once into a HIP graph (per residual,
 * output and done-flag address) and replayed; off: launched kernel by kernel.
 * Both give the same bits. */
int nbx_bp_mg_use_graph(nbx_bp* bp, int on);
/* order and lambda_max estimate of a level; returns the number of levels */
int nbx_bp_mg_info(const nbx_bp* bp, int level, int* N, double* lmax);
/* z = M r, one V-cycle (device vectors), synchronous */
int nbx_bp_mg_apply(nbx_bp* bp, const double* r, double* z);
/* Learned multigrid (no counterpart in the reference): with on = 1 the fine
 * level ends every V-cycle with one application of the network attached by
 * nbx_bp_attach_net to the residual the cycle leaves there:
 * rf = res - A d after post-smoothing, x += mask(gs(w invDiag (rf + s f(rf))))
 * (the network's element output, assembled).  N = 7, one field.  Networks
 * for this step are trained on pairs from nbx_net_generate_mg_pairs. */
int nbx_bp_mg_set_learned(nbx_bp* bp, int on);
int nbx_bp_mg_get_learned(const nbx_bp* bp);

/* ======================= learned preconditioner ==========================
 * A per-element network over the 8^3 GLL nodes (N = 7, one field): inputs
 * the residual (scaled by the element's max |r|) and the 7 geometric factors
 * (each divided by its RMS over the mesh); three layers (3x3x3 convolution of
 * the residual plus a pointwise map of the factors -> 32 channels, ReLU;
 * 32 -> 32 pointwise, ReLU; 3x3x3 convolution 32 -> 1); bf16 weights, FP32
 * arithmetic.  Output z_e = invDiag (r + s f), assembled by gather-scatter
 * with weights 1/multiplicity.  The zero-initialised output layer makes the
 * untrained preconditioner Jacobi.  Used by nbx_bp_pcg with
 * precon = NBX_PRECON_LEARNED inside flexible PCG; after a stall (|r|^2 not
 * below ratio x its value window iterations earlier) the solve continues with
 * Jacobi.  No counterpart in the reference (new feature of this build). */
typedef struct nbx_net nbx_net;
int nbx_net_num_params(void);
nbx_net* nbx_net_create(uint64_t seed, void* stream);
void nbx_net_free(nbx_net* net);  /* detaches it from every BP handle it is attached to */
int nbx_net_get_params(const nbx_net* net, float* host);      /* FP32 master weights */
int nbx_net_set_params(nbx_net* net, const float* host);
int nbx_net_save(const nbx_net* net, const char* path);        /* header, seed, params, checksum */
nbx_net* nbx_net_load(const char* path, void* stream);
int nbx_bp_attach_net(nbx_bp* bp, nbx_net* net);  /* the handle uses the network until either is freed */
/* inference used by the solve and nbx_net_apply: 0 (default) FP32 on the
 * vector ALUs, bit-exact with the oracle; 1 bf16 operands on the matrix
 * cores with FP32 accumulation (training always uses the FP32 forward) */
int nbx_net_set_inference(nbx_net* net, int mode);
int nbx_bp_set_fallback(nbx_bp* bp, int stall_window, double stall_ratio);
int nbx_bp_update_jacobi(nbx_bp* bp);                          /* invDiagA of the handle */
/* element-local output (weighted: times 1/multiplicity) for r (device) */
int nbx_net_apply(nbx_bp* bp, const double* r, double* zl, int weighted);
/* training pairs (r_k, x* - x_k) from Jacobi PCG snapshots; returns the count */
int nbx_net_generate_pairs(nbx_net* net, nbx_bp* bp, int nseeds, uint64_t seed0, const int* snapshots, int nsnap,
                           double tol_exact);
int nbx_net_set_pairs(nbx_net* net, nbx_bp* bp, int npairs, const float* R, const float* E);
/* training pairs for the learned multigrid step: for nseeds seeded right-hand
 * sides b, x* from the multigrid PCG to tol_exact and its iterates x_k,
 * k = kfirst .. kfirst+kcount-1; one V-cycle z = M (b - A x_k) leaves the
 * pair (r' = b - A x_k - A z, e' = x* - x_k - z).  The handle needs precon
 * MG (nbx_bp_mg_setup); returns the count */
int nbx_net_generate_mg_pairs(nbx_net* net, nbx_bp* bp, int nseeds, uint64_t seed0, int kfirst, int kcount,
                              double tol_exact);
/* host copies (R, E may be NULL) of the stored pairs first .. first+count-1,
 * local-vector length each */
int nbx_net_copy_pairs(const nbx_net* net, int first, int count, float* R, float* E);
/* loss sum_samples sum_n (z - e)^2 / sum_pair e^2 and its parameter gradient
 * over (pair, element) samples (host int pairs), chunk samples per workgroup.
 * G (device, local vector, may be NULL): when given, the loss is external and
 * G[e*Np+n] is its derivative with respect to the unweighted element output
 * invDiag (r + s f(r)); the returned loss is then 0 */
int nbx_net_gradient(nbx_net* net, nbx_bp* bp, const int32_t* samples, int nsamples, int chunk, const double* G,
                     float* grad, double* loss);
/* Adam on the L2 error loss above, batch random (pair, element) samples a step */
int nbx_net_train(nbx_net* net, nbx_bp* bp, int steps, int batch, float lr, uint64_t seed, double* loss_hist);
/* Adam on an energy-norm loss, one random pair a step, all its elements;
 * z = mask(gs(w invDiag (r + s f(r)))):
 *   mode 0: L = (z - e)^T A (z - e) / e^T A e
 *   mode 1: L = 1 - <r,z>^2 / (e^T A e <z,Az>)  (scale-invariant: the error
 *           left by the best multiple of z) */
int nbx_net_train_energy(nbx_net* net, nbx_bp* bp, int steps, float lr, uint64_t seed, int mode,
                         double* loss_hist);

/* Phase timers of nekBone's PROFILING breakdown (nekBone/main.cpp:181-224,
 * timer tags of nekBone/solve.cpp:71-109).  With profiling on, every PCG
 * solve records events at its phase boundaries and accumulates device time
 * per phase: seconds[0] local Ax (with the fused p = z + beta p and the pAp
 * shares), [1] gs (gather-scatter + mask), [2] updatePCG (x, r update, and
 * with COPY/JACOBI the preconditioner and the dots' tile partials, fused into
 * one pass), [3] dot (the reductions' finishing passes; with MG/LEARNED also
 * the dots), [4] preco (the V-cycle or the network with its assembly).
 * The events add a little time to each phase.  reset != 0 zeroes the sums. */
int nbx_bp_set_profiling(nbx_bp* bp, int on);
int nbx_bp_profile(nbx_bp* bp, double* seconds, int reset);

/* per-iteration history of the last solve (host outputs, length >= iterations+1):
 * rdotr and alpha per iteration; returns the count written */
int nbx_bp_history(const nbx_bp* bp, double* rdotr, double* alpha, int capacity);

#ifdef __cplusplus
}
#endif
#endif
