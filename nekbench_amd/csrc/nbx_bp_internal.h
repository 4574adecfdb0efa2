// Internal state of the BP driver, shared by nbx_bp.hip and nbx_learned.hip.
#pragma once
#include "nbx_blas.h"
#include <vector>

#include <cstddef>

struct nbx_net;
struct nbx_mg;

namespace nbx {
// the leading members are PcgFlags (nbx_internal.h), read by operator kernels
struct PcgScalars {
  int done, pad_;
  double beta;
  double rdotz1, rdotz2, rdotr, rdotr0, TOL, tol2, alpha, pAp, zdotAp, abs_floor;
  int iter, iterations, converged, breakdown, maxit, fixed, flexible;
  int learned, fallback, stall_window, fallback_iter;  // learned preconditioner state
  double stall_ratio;
};
static_assert(offsetof(PcgScalars, done) == offsetof(PcgFlags, done) &&
                  offsetof(PcgScalars, beta) == offsetof(PcgFlags, beta),
              "PcgScalars must start with PcgFlags");
}  // namespace nbx

struct nbx_bp {
  nbx_bp_options opt{};
  int N = 0, Nq = 0, Np = 0, Nf = 1;
  int32_t E = 0, Nlocal = 0, fieldOffset = 0, Nmasked = 0;
  int64_t Nall = 0;
  // singular all-Neumann problem (lambda1 = 0, setup.cpp:210-214) without a
  // Dirichlet mask: r and x are projected to zero mean around the solve
  bool allNeumann = false;
  hipStream_t stream = nullptr;
  nbx_ogs* ogs = nullptr;
  double *D = nullptr, *ggeo = nullptr, *lambda = nullptr, *invDiagA = nullptr, *rhs = nullptr;
  double *p = nullptr, *z = nullptr, *Ap = nullptr, *Ax = nullptr;
  double* elem_partials = nullptr;  // per-element shares of pAp (fused schedule)
  // low-precision operator operands (built on first use)
  float* ggeo32 = nullptr;
  float* D32 = nullptr;
  float* lambda32 = nullptr;
  uint16_t* ggeo16 = nullptr;
  uint16_t* lambda16 = nullptr;
  double *b = nullptr, *d = nullptr;   // refinement: right-hand side copy, correction
  nbx_prec op_prec = NBX_F64;         // operator precision of the PCG being run
  int32_t* maskIds = nullptr;
  nbx::ReduceWork red;
  nbx::PcgScalars* S = nullptr;
  nbx::PcgScalars* hS = nullptr;  // pinned host copy
  double *hist_rdotr = nullptr, *hist_alpha = nullptr;
  int hist_cap = 0;
  int last_iters = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // learned preconditioner (attached network, features, per-element output)
  nbx_net* net = nullptr;
  nbx_mg* mg = nullptr;     // p-multigrid preconditioner (levels own their handles)
  float* feat = nullptr;    // normalised geometric factors [E][7][Np]
  uint16_t* feat16 = nullptr;  // the same as bf16 (matrix-core inference)
  double* zl = nullptr;     // element-local preconditioner output
  int stall_window = 20;    // iterations compared for the stagnation test
  double stall_ratio = 0.9; // rdotr(it) > ratio * rdotr(it - window) -> Jacobi from then on
  std::vector<void*> owned;
  // phase timers of nekBone's PROFILING breakdown (nekBone/main.cpp:181-224):
  // events recorded at phase boundaries, each interval credited to the phase
  // that ends there; accumulated at every host wait of the solve
  bool prof = false;
  std::vector<hipEvent_t> prof_ev;
  std::vector<int> prof_tag;
  int prof_used = 0;
  double prof_t[5] = {0, 0, 0, 0, 0};  // Ax, gs, updatePCG, dot, preco

  int alloc(double** ptr, int64_t n)
  {
    NBX_HIP(hipMalloc((void**)ptr, sizeof(double) * (size_t)std::max<int64_t>(n, 1)));
    NBX_HIP(hipMemset(*ptr, 0, sizeof(double) * (size_t)std::max<int64_t>(n, 1)));
    owned.push_back(*ptr);
    return NBX_SUCCESS;
  }
};


namespace nbx {
enum ProfPhase { kProfAx = 0, kProfGs = 1, kProfUpdate = 2, kProfDot = 3, kProfPreco = 4, kProfNone = -1 };
// record a phase boundary (no-op unless profiling); the interval since the
// previous boundary is credited to `phase`
void prof_mark(nbx_bp* bp, int phase);
// accumulate the recorded intervals (the stream must have reached the last mark)
int prof_collect(nbx_bp* bp);
}  // namespace nbx

// FP64 PCG on the handle's operator with an explicit preconditioner choice
// (used to generate training data); x is updated in place, r becomes the
// residual.  tol <= 0 runs exactly maxit iterations.
int bp_solve_plain(nbx_bp* bp, nbx_precon precon, double* r, double* x, double tol, int maxit, int* iterations);
// Jacobi diagonal of the handle's operator with this lambda into invDiagA
int bp_update_jacobi(nbx_bp* bp, const double* lambda);
// enqueue one PCG iteration body (device-resident scalars; no host wait)
int bp_enqueue_iteration(nbx_bp* bp, const double* lambda, double* r, double* x);
// FP32 / bf16 copies of D, ggeo and lambda for the low-precision operator
int bp_prepare_precision(nbx_bp* bp, nbx_prec prec, const double* lambda);
// Aq = mask(gs(axhelm(q))) with the handle's FP64 operator (no dot)
int bp_apply(nbx_bp* bp, const double* lambda, const double* q, double* Aq, const int* skip);
// seeded right-hand side of the handle's problem (U[-1,1) per local node,
// stream 3 + field, masked, assembled) into out (device, Nall)
int bp_make_rhs(nbx_bp* bp, uint64_t seed, double* out);

namespace nbx {
// element-local learned preconditioner output w invDiag (r + s f(r)) of the
// attached network into out (Jacobi when *fallback)
int net_apply_learned(nbx_bp* bp, const double* r, double* out, const int* fallback, const int* skip);
// matrix-core inference only: the input residual is (double)(res - Ad)
int net_apply_learned_resid(nbx_bp* bp, const float* res, const float* Ad, double* out, const int* skip);
// forget a handle that is being freed (the network keeps a list of the
// handles it is attached to and clears their pointer when it is freed first)
void net_detach(nbx_net* net, nbx_bp* bp);
// inference kernel of a network (0 FP32, 1 matrix cores; -1 for none)
int net_inference_mode(const nbx_net* net);
// one multigrid V-cycle z = M r on the handle's stream (kernels skip when *skip)
int mg_apply(nbx_bp* bp, const double* r, double* z, const int* skip);
// the same V-cycle launched kernel by kernel (no graph capture or replay)
int mg_apply_direct(nbx_bp* bp, const double* r, double* z);
void mg_release(nbx_mg* mg);
}  // namespace nbx
