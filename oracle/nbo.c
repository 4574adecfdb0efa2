/*
 * nbo.c -- CPU oracle for the nekBench solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU
 * baseline leg of bench.py as the checker; never linked by the product.
 *
 * Build: see oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).
 *
 * Pinning (what checks each function besides retyping it):
 *   nbo_gll / nbo_dmatrix   closed forms: sum_i w_i = 2, exact quadrature of
 *                           degree <= 2N-1, D * p = p' for degree <= N,
 *                           rows of D sum to zero (tests/test_oracle_basis.py)
 *   nbo_geometric_factors   affine elements give diagonal G = JW/h^2, sum GWJ
 *                           = box volume for any perturbation (tests/test_oracle_mesh.py)
 *   nbo_global_ids          multiplicity histogram {1,2,4,8} of a box, counts
 *                           of unique nodes, degree round trip of gs/main.cpp
 *   nbo_ogs_setup/gs_many   brute-force grouping by sorting, gslib multiplicity
 *                           asserts (gs_test.c:88-105), integer round trip
 *   nbo_axhelm              the compiled reference axhelm.c / axhelmReference.cpp
 *                           (oracle/_ref) and golden vectors generated from them,
 *                           A*1 = 0, symmetry, exact energy of polynomials
 *   nbo_jacobi_local        diagonal of the dense element matrix built from
 *                           nbo_axhelm applied to unit vectors
 *   nbo_tree_*              exact integer-valued sums, brute-force numpy sums
 *   nbo_axhelm_bf16mma /    within operand rounding of nbo_axhelm (FP64) for
 *   nbo_axhelm_f16mma       BK and non-BK, constant vector near the null space;
 *                           the fp16 form several times closer than the bf16
 *                           form (tests/test_oracle_axhelm.py)
 *   nbo_f16_round_array     numpy's float16 conversion: random values, exact
 *                           ties (normal and subnormal), overflow edge, zeros
 */
#include "nbo.h"
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GWJID 0
#define G00ID 1
#define G01ID 2
#define G02ID 3
#define G11ID 4
#define G12ID 5
#define G22ID 6
#define NGGEO 7

/* ------------------------------------------------------------------------ */
uint64_t nbo_splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

double nbo_urand(uint64_t seed, uint64_t stream, uint64_t i)
{
  const uint64_t key = nbo_splitmix64(seed ^ nbo_splitmix64(stream));
  const uint64_t bits = nbo_splitmix64(key + i);
  return (double)(bits >> 11) * 0x1.0p-52 - 1.0;
}

void nbo_urand_fill(uint64_t seed, uint64_t stream, int64_t n, double* out)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = nbo_urand(seed, stream, (uint64_t)i);
}

/* ------------------------------------------------------------------------ */
/* Legendre P_N and P_N' by the three-term recurrence. */
static void legendre(int N, double x, double* P, double* dP)
{
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  if (N == 0) { *P = 1.0; *dP = 0.0; return; }
  for (int n = 2; n <= N; ++n) {
    const double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
    const double d2 = d0 + (2.0 * n - 1.0) * p1;  /* P_n' = P_{n-2}' + (2n-1) P_{n-1} */
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  *P = p1; *dP = d1;
}

/* GLL nodes = {-1, roots of P_N', +1} ascending; w_i = 2/(N(N+1) P_N(z_i)^2).
 * Same nodes and weights as meshJacobiGL (core/meshBasis.cpp:140-176), which
 * gets the interior nodes as Gauss-Jacobi(1,1) points and the weights by mass
 * lumping; here they come from Newton on P_N' (Chebyshev-Lobatto start). */
void nbo_gll(int N, double* z, double* w)
{
  const double PI = 3.14159265358979323846;
  z[0] = -1.0; z[N] = 1.0;
  for (int i = 1; 2 * i < N + 1; ++i) {
    if (2 * i == N) { z[i] = 0.0; continue; }
    double x = -cos(PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(N, x, &P, &dP);
      const double d2P = (2.0 * x * dP - N * (N + 1.0) * P) / (1.0 - x * x);
      const double dx = dP / d2P;
      x -= dx;
      if (fabs(dx) < 1e-17) break;
    }
    z[i] = x;
    z[N - i] = -x;
  }
  for (int i = 0; i <= N; ++i) {
    double P, dP;
    legendre(N, z[i], &P, &dP);
    w[i] = 2.0 / (N * (N + 1.0) * P * P);
  }
}

/* D_ij = l_j'(z_i) (core/meshBasis.cpp:775-790 computes Vr V^-1; closed form
 * used here): P_N(z_i)/(P_N(z_j)(z_i-z_j)), D_00 = -N(N+1)/4, D_NN = +N(N+1)/4. */
void nbo_dmatrix(int N, const double* z, double* D)
{
  const int Nq = N + 1;
  double* P = (double*)malloc(sizeof(double) * Nq);
  for (int i = 0; i < Nq; ++i) { double dP; legendre(N, z[i], &P[i], &dP); }
  for (int i = 0; i < Nq; ++i)
    for (int j = 0; j < Nq; ++j) {
      double d = 0.0;
      if (i != j) d = P[i] / (P[j] * (z[i] - z[j]));
      else if (i == 0) d = -N * (N + 1.0) / 4.0;
      else if (i == N) d = N * (N + 1.0) / 4.0;
      D[i * Nq + j] = d;
    }
  free(P);
}

/* ------------------------------------------------------------------------ */
int64_t nbo_box_nelements(const nbo_box* b) { return (int64_t)b->NX * b->NY * b->NZ; }

/* z-plane coordinates: uniform, or geometric with first layer hx/zaspect */
static void zplanes(const nbo_box* b, double* zp)
{
  const int NZ = b->NZ;
  if (b->zaspect <= 0.0) {
    for (int k = 0; k <= NZ; ++k) zp[k] = b->z0 + (b->z1 - b->z0) * k / NZ;
    return;
  }
  const double Lz = b->z1 - b->z0;
  const double hx = (b->x1 - b->x0) / b->NX;
  const double dz0 = hx / b->zaspect;
  /* growth g > 1 with dz0 * (g^NZ - 1)/(g - 1) = Lz, bisection */
  double lo = 1.0 + 1e-12, hi = 4.0;
  for (int it = 0; it < 200; ++it) {
    const double g = 0.5 * (lo + hi);
    const double s = dz0 * (pow(g, NZ) - 1.0) / (g - 1.0);
    if (s > Lz) hi = g; else lo = g;
  }
  const double g = 0.5 * (lo + hi);
  for (int k = 0; k <= NZ; ++k) zp[k] = b->z0 + Lz * (pow(g, k) - 1.0) / (pow(g, NZ) - 1.0);
  zp[NZ] = b->z1;
}

/* Vertices of element (ex,ey,ez) in the reference order
 * (core/meshSetup.cpp:2378-2410): 0:(i,j,k) 1:(i+1,j,k) 2:(i+1,j+1,k)
 * 3:(i,j+1,k), 4-7 the same at k+1.  Interior box vertices move by
 * perturb * (local width) * U[-1,1) per axis, stream 1 of the seed. */
void nbo_box_vertices(const nbo_box* b, double* EX, double* EY, double* EZ, int64_t* EToV)
{
  const int NX = b->NX, NY = b->NY, NZ = b->NZ;
  const int64_t VX = NX + 1, VY = NY + 1;
  double* zp = (double*)malloc(sizeof(double) * (NZ + 1));
  zplanes(b, zp);
  const double hx = (b->x1 - b->x0) / NX, hy = (b->y1 - b->y0) / NY;
  const int off[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
  const int64_t E = nbo_box_nelements(b);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    const int ex = (int)(e % NX), ey = (int)((e / NX) % NY), ez = (int)(e / ((int64_t)NX * NY));
    for (int v = 0; v < 8; ++v) {
      const int I = ex + off[v][0], J = ey + off[v][1], K = ez + off[v][2];
      double x = b->x0 + hx * I, y = b->y0 + hy * J, z = zp[K];
      if (b->perturb != 0.0 && I > 0 && I < NX && J > 0 && J < NY && K > 0 && K < NZ) {
        const uint64_t vid = (uint64_t)(I + VX * (J + VY * (int64_t)K));
        const double hz = fmin(zp[K] - zp[K - 1], zp[K + 1] - zp[K]);
        x += b->perturb * hx * nbo_urand(b->seed, 1, 3 * vid + 0);
        y += b->perturb * hy * nbo_urand(b->seed, 1, 3 * vid + 1);
        z += b->perturb * hz * nbo_urand(b->seed, 1, 3 * vid + 2);
      }
      EX[e * 8 + v] = x; EY[e * 8 + v] = y; EZ[e * 8 + v] = z;
      int64_t gv;
      if (b->periodic)
        gv = (I % NX) + (int64_t)NX * ((J % NY) + (int64_t)NY * (K % NZ));
      else
        gv = I + VX * (J + VY * (int64_t)K);
      if (EToV) EToV[e * 8 + v] = gv;
    }
  }
  free(zp);
}

/* core/meshSetup.cpp:2156-2233, same expression and evaluation order */
void nbo_physical_nodes(int N, int64_t E, const double* EX, const double* EY, const double* EZ,
                        const double* zr, double* x, double* y, double* z)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    const double* xe = EX + e * 8; const double* ye = EY + e * 8; const double* ze = EZ + e * 8;
    for (int n = 0; n < Np; ++n) {
      const double rn = zr[n % Nq], sn = zr[(n / Nq) % Nq], tn = zr[n / (Nq * Nq)];
      const int64_t c = e * Np + n;
#define TRILIN(v) \
      (+0.125 * (1 - rn) * (1 - sn) * (1 - tn) * v[0] \
       + 0.125 * (1 + rn) * (1 - sn) * (1 - tn) * v[1] \
       + 0.125 * (1 + rn) * (1 + sn) * (1 - tn) * v[2] \
       + 0.125 * (1 - rn) * (1 + sn) * (1 - tn) * v[3] \
       + 0.125 * (1 - rn) * (1 - sn) * (1 + tn) * v[4] \
       + 0.125 * (1 + rn) * (1 - sn) * (1 + tn) * v[5] \
       + 0.125 * (1 + rn) * (1 + sn) * (1 + tn) * v[6] \
       + 0.125 * (1 - rn) * (1 + sn) * (1 + tn) * v[7])
      x[c] = TRILIN(xe);
      y[c] = TRILIN(ye);
      z[c] = TRILIN(ze);
#undef TRILIN
    }
  }
}

/* core/meshSetup.cpp:266-360: nodal derivatives through D, then the inverse
 * Jacobian, JW = J w_i w_j w_k, G_ab = JW (grad r_a . grad r_b). */
double nbo_geometric_factors(int N, int64_t E, const double* x, const double* y, const double* z,
                             const double* D, const double* w, double* ggeo)
{
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  double minJ = 1e300;
#pragma omp parallel for schedule(static) reduction(min : minJ)
  for (int64_t e = 0; e < E; ++e) {
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq * Nq;
          double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr = 0, zs = 0, zt = 0;
          for (int m = 0; m < Nq; ++m) {
            const int64_t idr = e * Np + k * Nq * Nq + j * Nq + m;
            const int64_t ids = e * Np + k * Nq * Nq + m * Nq + i;
            const int64_t idt = e * Np + m * Nq * Nq + j * Nq + i;
            xr += D[i * Nq + m] * x[idr]; xs += D[j * Nq + m] * x[ids]; xt += D[k * Nq + m] * x[idt];
            yr += D[i * Nq + m] * y[idr]; ys += D[j * Nq + m] * y[ids]; yt += D[k * Nq + m] * y[idt];
            zr += D[i * Nq + m] * z[idr]; zs += D[j * Nq + m] * z[ids]; zt += D[k * Nq + m] * z[idt];
          }
          const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt);
          if (J < minJ) minJ = J;
          const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
          const double sx = -(yr * zt - zr * yt) / J, sy = (xr * zt - zr * xt) / J, sz = -(xr * yt - yr * xt) / J;
          const double tx = (yr * zs - zr * ys) / J, ty = -(xr * zs - zr * xs) / J, tz = (xr * ys - yr * xs) / J;
          const double JW = J * w[i] * w[j] * w[k];
          double* g = ggeo + e * (int64_t)NGGEO * Np + n;
          g[Np * G00ID] = JW * (rx * rx + ry * ry + rz * rz);
          g[Np * G01ID] = JW * (rx * sx + ry * sy + rz * sz);
          g[Np * G02ID] = JW * (rx * tx + ry * ty + rz * tz);
          g[Np * G11ID] = JW * (sx * sx + sy * sy + sz * sz);
          g[Np * G12ID] = JW * (sx * tx + sy * ty + sz * tz);
          g[Np * G22ID] = JW * (tx * tx + ty * ty + tz * tz);
          g[Np * GWJID] = JW;
        }
  }
  return minJ;
}

/* Fixed point of the min-label propagation of meshParallelConnectNodes
 * (core/meshSetup.cpp:1638-1735) on one rank: every node of a connected group
 * ends with the smallest initial label in the group.  Initial labels are
 * EToV+1 at element vertices and 1 + id + Nnodes elsewhere, so a vertex group
 * carries its vertex number + 1 and any other group 1 + Nnodes + (smallest
 * local index in the group).  Groups are identified here by the structured
 * global GLL index of each node. */
void nbo_global_ids(const nbo_box* b, int mask_boundary, int64_t* ids)
{
  const int N = b->N, Nq = N + 1, Np = Nq * Nq * Nq;
  const int NX = b->NX, NY = b->NY, NZ = b->NZ;
  const int64_t E = nbo_box_nelements(b);
  const int64_t GX = b->periodic ? (int64_t)NX * N : (int64_t)NX * N + 1;
  const int64_t GY = b->periodic ? (int64_t)NY * N : (int64_t)NY * N + 1;
  const int64_t GZ = b->periodic ? (int64_t)NZ * N : (int64_t)NZ * N + 1;
  const int64_t Nnodes = b->periodic ? (int64_t)NX * NY * NZ : (int64_t)(NX + 1) * (NY + 1) * (NZ + 1);
  int64_t* first = (int64_t*)malloc(sizeof(int64_t) * GX * GY * GZ);
  for (int64_t g = 0; g < GX * GY * GZ; ++g) first[g] = -1;
  for (int64_t e = 0; e < E; ++e) {
    const int ex = (int)(e % NX), ey = (int)((e / NX) % NY), ez = (int)(e / ((int64_t)NX * NY));
    for (int n = 0; n < Np; ++n) {
      const int i = n % Nq, j = (n / Nq) % Nq, k = n / (Nq * Nq);
      int64_t I = (int64_t)ex * N + i, J = (int64_t)ey * N + j, K = (int64_t)ez * N + k;
      const int onbnd = (I == 0 || J == 0 || K == 0 || I == (int64_t)NX * N || J == (int64_t)NY * N ||
                         K == (int64_t)NZ * N);
      if (b->periodic) { I %= GX; J %= GY; K %= GZ; }
      const int64_t g = I + GX * (J + GY * K);
      const int64_t id = e * Np + n;
      if (first[g] < 0) first[g] = id;
      const int isvert = (i == 0 || i == N) && (j == 0 || j == N) && (k == 0 || k == N);
      int64_t label;
      if (isvert) {
        const int64_t vi = ex + (i == N), vj = ey + (j == N), vk = ez + (k == N);
        label = b->periodic ? (vi % NX) + (int64_t)NX * ((vj % NY) + (int64_t)NY * (vk % NZ))
                            : vi + (int64_t)(NX + 1) * (vj + (int64_t)(NY + 1) * vk);
        label += 1;
      } else {
        label = 1 + first[g] + Nnodes;
      }
      ids[id] = (mask_boundary && !b->periodic && onbnd) ? 0 : label;
    }
  }
  free(first);
}

/* ------------------------------------------------------------------------ */
typedef struct { int32_t localId; int64_t baseId; int32_t newId; int owned; } pnode;

static int cmp_base(const void* a, const void* b)
{
  const pnode* fa = (const pnode*)a; const pnode* fb = (const pnode*)b;
  const int64_t A = llabs(fa->baseId), B = llabs(fb->baseId);
  if (A < B) return -1;
  if (A > B) return 1;
  if (fa->localId < fb->localId) return -1;
  if (fa->localId > fb->localId) return 1;
  return 0;
}
static int cmp_local(const void* a, const void* b)
{
  const pnode* fa = (const pnode*)a; const pnode* fb = (const pnode*)b;
  return (fa->localId > fb->localId) - (fa->localId < fb->localId);
}

/* Local part of ogsSetup (3rdParty/ogs/src/ogsSetup.cpp:137-212) for one rank
 * (every node is local): sort by (|id|, local id), number groups, re-sort by
 * local id, give groups their final number in order of first appearance,
 * members ascending.  invDegree as in :325-352 (1 for excluded nodes). */
int nbo_ogs_setup(int32_t N, const int64_t* ids, nbo_ogs* o)
{
  memset(o, 0, sizeof(*o));
  o->N = N;
  int32_t Nlocal = 0;
  for (int32_t i = 0; i < N; ++i) if (ids[i] != 0) ++Nlocal;
  o->Nlocal = Nlocal;
  o->invDegree = (double*)malloc(sizeof(double) * (N > 0 ? N : 1));
  for (int32_t i = 0; i < N; ++i) o->invDegree[i] = 1.0;
  o->offsets = (int32_t*)calloc((size_t)Nlocal + 1, sizeof(int32_t));
  o->gatherIds = (int32_t*)calloc(Nlocal > 0 ? Nlocal : 1, sizeof(int32_t));
  if (!Nlocal) return 0;
  pnode* nodes = (pnode*)calloc(Nlocal, sizeof(pnode));
  int32_t cnt = 0;
  for (int32_t i = 0; i < N; ++i)
    if (ids[i] != 0) { nodes[cnt].localId = i; nodes[cnt].baseId = ids[i]; cnt++; }
  qsort(nodes, Nlocal, sizeof(pnode), cmp_base);
  int32_t NlocalGather = 0;
  nodes[0].newId = 0; nodes[0].owned = 1;
  for (int32_t i = 1; i < Nlocal; ++i) {
    int s = 0;
    if (nodes[i].baseId != nodes[i - 1].baseId) { NlocalGather++; s = 1; }
    nodes[i].newId = NlocalGather; nodes[i].owned = s;
  }
  NlocalGather++;
  qsort(nodes, Nlocal, sizeof(pnode), cmp_local);
  int32_t* counts = (int32_t*)calloc(NlocalGather, sizeof(int32_t));
  int32_t* map = (int32_t*)calloc(NlocalGather, sizeof(int32_t));
  cnt = 0;
  for (int32_t i = 0; i < Nlocal; ++i) {
    const int32_t newId = nodes[i].newId;
    if (nodes[i].owned) map[newId] = cnt++;
    nodes[i].newId = map[newId];
    counts[map[newId]]++;
  }
  free(map);
  int32_t* offsets = (int32_t*)realloc(o->offsets, sizeof(int32_t) * ((size_t)NlocalGather + 1));
  offsets[0] = 0;
  for (int32_t g = 0; g < NlocalGather; ++g) { offsets[g + 1] = offsets[g] + counts[g]; counts[g] = 0; }
  for (int32_t i = 0; i < Nlocal; ++i) {
    const int32_t g = nodes[i].newId;
    o->gatherIds[offsets[g] + counts[g]] = nodes[i].localId;
    counts[g]++;
  }
  free(counts);
  free(nodes);
  o->offsets = offsets;
  o->NlocalGather = NlocalGather;
  for (int32_t g = 0; g < NlocalGather; ++g) {
    const double deg = (double)(offsets[g + 1] - offsets[g]);
    for (int32_t n = offsets[g]; n < offsets[g + 1]; ++n) o->invDegree[o->gatherIds[n]] = 1.0 / deg;
  }
  return 0;
}

void nbo_ogs_free(nbo_ogs* o)
{
  free(o->offsets); free(o->gatherIds); free(o->invDegree);
  memset(o, 0, sizeof(*o));
}

/* gatherScatterMany (3rdParty/ogs/okl/gatherScatterMany.okl:58-86 add,
 * :148-176 mul, :300-330 min, :455-485 max): per group and field, add/mul
 * start from 0/1 and combine every member in ascending order; min/max start
 * from the first member and compare the rest with (b < acc) ? b : acc.  The
 * result goes back to every member; singletons are skipped.  Groups are
 * disjoint, so they run in parallel with the same result. */
#define GS_LOOP(T, FIRST, INIT, COMB)                                             \
  do {                                                                            \
    T* v = (T*)q;                                                                 \
    for (int f = 0; f < k; ++f)                                                   \
      _Pragma("omp parallel for schedule(static)")                                \
      for (int32_t g = 0; g < o->NlocalGather; ++g) {                            \
        const int32_t s = o->offsets[g], e = o->offsets[g + 1];                   \
        if (s + 1 == e) continue;                                                 \
        T acc = FIRST ? v[o->gatherIds[s] + (int64_t)f * stride] : (T)(INIT);     \
        for (int32_t n = s + (FIRST ? 1 : 0); n < e; ++n) {                       \
          const T b = v[o->gatherIds[n] + (int64_t)f * stride];                   \
          COMB;                                                                   \
        }                                                                         \
        for (int32_t n = s; n < e; ++n) v[o->gatherIds[n] + (int64_t)f * stride] = acc; \
      }                                                                           \
  } while (0)

#define GS_OPS(T)                                                                 \
  switch (op) {                                                                   \
    case NBO_ADD: GS_LOOP(T, 0, 0, acc = acc + b); break;                        \
    case NBO_MUL: GS_LOOP(T, 0, 1, acc = acc * b); break;                        \
    case NBO_MIN: GS_LOOP(T, 1, 0, acc = (b < acc) ? b : acc); break;            \
    case NBO_MAX: GS_LOOP(T, 1, 0, acc = (b > acc) ? b : acc); break;            \
    default: return -1;                                                           \
  }

int nbo_gs_many(const nbo_ogs* o, void* q, int k, int32_t stride, int type, int op)
{
  switch (type) {
    case NBO_DOUBLE: GS_OPS(double); break;
    case NBO_FLOAT: GS_OPS(float); break;
    case NBO_INT: GS_OPS(int32_t); break;
    case NBO_LONG: GS_OPS(long long); break;
    default: return -1;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* One element at a compile-time Nq (one copy per order, below; kept outside
 * the OpenMP region so the order stays a constant inside it), with D also
 * held transposed, so that every loop over i reads contiguous data and the
 * compiler vectorises across the nodes of a row; each node's fma chains are
 * the ones written here, in this order, whatever the vector width. */
static inline __attribute__((always_inline)) void axhelm_elem(
    const int Nq, const int64_t e, const int Nf, const int bk, const int32_t offset, const double* restrict ggeo,
    const double* restrict D, const double* restrict DT, const double* restrict lambda, const double* restrict q,
    double* restrict Aq, double* restrict s_q, double* restrict s_r, double* restrict s_s, double* restrict s_t)
{
  const int Nq2 = Nq * Nq, Np = Nq2 * Nq;
  const double* restrict g = ggeo + e * (int64_t)NGGEO * Np;
  for (int f = 0; f < Nf; ++f) {
    const double* restrict qe = q + e * Np + (int64_t)f * offset;
    double* restrict Aqe = Aq + e * Np + (int64_t)f * offset;
    for (int n = 0; n < Np; ++n) s_q[n] = qe[n];
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq2;
          double qr = 0.0, qs = 0.0, qt = 0.0;
#pragma GCC unroll 16
          for (int m = 0; m < Nq; ++m) {
            qr = fma(DT[m * Nq + i], s_q[m + j * Nq + k * Nq2], qr);
            qs = fma(D[j * Nq + m], s_q[i + m * Nq + k * Nq2], qs);
            qt = fma(D[k * Nq + m], s_q[i + j * Nq + m * Nq2], qt);
          }
          const double G00 = g[G00ID * Np + n], G01 = g[G01ID * Np + n], G02 = g[G02ID * Np + n];
          const double G11 = g[G11ID * Np + n], G12 = g[G12ID * Np + n], G22 = g[G22ID * Np + n];
          double Gr = fma(G02, qt, fma(G01, qs, G00 * qr));
          double Gs = fma(G12, qt, fma(G11, qs, G01 * qr));
          double Gt = fma(G22, qt, fma(G12, qs, G02 * qr));
          if (!bk) {
            const double l0 = lambda[e * Np + n];
            Gr = l0 * Gr; Gs = l0 * Gs; Gt = l0 * Gt;
          }
          s_r[n] = Gr; s_s[n] = Gs; s_t[n] = Gt;
        }
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq2;
          double ar = 0.0, as = 0.0, at = 0.0;
#pragma GCC unroll 16
          for (int m = 0; m < Nq; ++m) {
            ar = fma(D[m * Nq + i], s_r[m + j * Nq + k * Nq2], ar);
            as = fma(D[m * Nq + j], s_s[i + m * Nq + k * Nq2], as);
            at = fma(D[m * Nq + k], s_t[i + j * Nq + m * Nq2], at);
          }
          double res = (ar + as) + at;
          if (!bk) {
            const double gl = g[GWJID * Np + n] * lambda[e * Np + n + offset];
            res = fma(gl, s_q[n], res);
          }
          Aqe[n] = res;
        }
  }
}

typedef void (*axhelm_elem_fn)(int64_t, int, int, int32_t, const double*, const double*, const double*,
                               const double*, const double*, double*, double*, double*, double*, double*);
#define AXHELM_NQ(NQ)                                                                                  \
  static void axhelm_elem##NQ(int64_t e, int Nf, int bk, int32_t offset, const double* ggeo,           \
                              const double* D, const double* DT, const double* lambda, const double* q, \
                              double* Aq, double* s_q, double* s_r, double* s_s, double* s_t)            \
  {                                                                                                    \
    if (bk) axhelm_elem(NQ, e, Nf, 1, offset, ggeo, D, DT, lambda, q, Aq, s_q, s_r, s_s, s_t);          \
    else axhelm_elem(NQ, e, Nf, 0, offset, ggeo, D, DT, lambda, q, Aq, s_q, s_r, s_s, s_t);             \
  }
AXHELM_NQ(2) AXHELM_NQ(3) AXHELM_NQ(4) AXHELM_NQ(5) AXHELM_NQ(6) AXHELM_NQ(7) AXHELM_NQ(8)
AXHELM_NQ(9) AXHELM_NQ(10) AXHELM_NQ(11) AXHELM_NQ(12) AXHELM_NQ(13) AXHELM_NQ(14) AXHELM_NQ(15)
AXHELM_NQ(16)

/* axhelm_v0 / axhelm_n3_v0 / axhelm_bk_v0 / axhelm_bk_n3_v0 and their Partial
 * forms (axhelm/kernel/CPU/axhelm.c:27-413; elementList as in
 * axhelm/kernel/VOLTA/axhelm.okl:158-165).  Per node:
 *   qr = sum_m D[i][m] q[k][j][m] (fma chain, m ascending), qs, qt alike
 *   G.q : Gqr = fma(G02,qt, fma(G01,qs, G00*qr)) ... ; non-BK scales by lambda0
 *   Aq  = ((Aqr + Aqs) + Aqt) [+ fma(GWJ*lambda1, q, .)]   */
void nbo_axhelm(int N, int Nf, int bk, int32_t Nelements, int32_t offset, const int32_t* elementList,
                const double* ggeo, const double* D, const double* lambda,
                const double* q, double* Aq)
{
  static const axhelm_elem_fn fns[15] = {axhelm_elem2,  axhelm_elem3,  axhelm_elem4,  axhelm_elem5,
                                         axhelm_elem6,  axhelm_elem7,  axhelm_elem8,  axhelm_elem9,
                                         axhelm_elem10, axhelm_elem11, axhelm_elem12, axhelm_elem13,
                                         axhelm_elem14, axhelm_elem15, axhelm_elem16};
  if (N < 1 || N > 15) return;
  const axhelm_elem_fn fn = fns[N - 1];
  const int Nq = N + 1, Np = Nq * Nq * Nq;
  double DT[16 * 16];
  for (int i = 0; i < Nq; ++i)
    for (int m = 0; m < Nq; ++m) DT[m * Nq + i] = D[i * Nq + m];
#pragma omp parallel
  {
    double* s_q = (double*)malloc(sizeof(double) * 4 * Np);
#pragma omp for schedule(static)
    for (int32_t ee = 0; ee < Nelements; ++ee) {
      const int64_t e = elementList ? elementList[ee] : ee;
      fn(e, Nf, bk, offset, ggeo, D, DT, lambda, q, Aq, s_q, s_q + Np, s_q + 2 * Np, s_q + 3 * Np);
    }
    free(s_q);
  }
}

void nbo_axhelm_f32(int N, int Nf, int bk, int32_t Nelements, int32_t offset,
                    const float* ggeo, const float* D, const float* lambda,
                    const float* q, float* Aq)
{
  const int Nq = N + 1, Nq2 = Nq * Nq, Np = Nq2 * Nq;
#pragma omp parallel
  {
    float* s_q = (float*)malloc(sizeof(float) * Np);
    float* s_r = (float*)malloc(sizeof(float) * Np);
    float* s_s = (float*)malloc(sizeof(float) * Np);
    float* s_t = (float*)malloc(sizeof(float) * Np);
#pragma omp for schedule(static)
    for (int32_t e = 0; e < Nelements; ++e) {
      const float* g = ggeo + (int64_t)e * NGGEO * Np;
      for (int f = 0; f < Nf; ++f) {
        const float* qe = q + (int64_t)e * Np + (int64_t)f * offset;
        float* Aqe = Aq + (int64_t)e * Np + (int64_t)f * offset;
        for (int n = 0; n < Np; ++n) s_q[n] = qe[n];
        for (int n = 0; n < Np; ++n) {
          const int i = n % Nq, j = (n / Nq) % Nq, k = n / Nq2;
          float qr = 0.f, qs = 0.f, qt = 0.f;
          for (int m = 0; m < Nq; ++m) {
            qr = fmaf(D[i * Nq + m], s_q[m + j * Nq + k * Nq2], qr);
            qs = fmaf(D[j * Nq + m], s_q[i + m * Nq + k * Nq2], qs);
            qt = fmaf(D[k * Nq + m], s_q[i + j * Nq + m * Nq2], qt);
          }
          float Gr = fmaf(g[G02ID * Np + n], qt, fmaf(g[G01ID * Np + n], qs, g[G00ID * Np + n] * qr));
          float Gs = fmaf(g[G12ID * Np + n], qt, fmaf(g[G11ID * Np + n], qs, g[G01ID * Np + n] * qr));
          float Gt = fmaf(g[G22ID * Np + n], qt, fmaf(g[G12ID * Np + n], qs, g[G02ID * Np + n] * qr));
          if (!bk) { const float l0 = lambda[(int64_t)e * Np + n]; Gr *= l0; Gs *= l0; Gt *= l0; }
          s_r[n] = Gr; s_s[n] = Gs; s_t[n] = Gt;
        }
        for (int n = 0; n < Np; ++n) {
          const int i = n % Nq, j = (n / Nq) % Nq, k = n / Nq2;
          float ar = 0.f, as = 0.f, at = 0.f;
          for (int m = 0; m < Nq; ++m) {
            ar = fmaf(D[m * Nq + i], s_r[m + j * Nq + k * Nq2], ar);
            as = fmaf(D[m * Nq + j], s_s[i + m * Nq + k * Nq2], as);
            at = fmaf(D[m * Nq + k], s_t[i + j * Nq + m * Nq2], at);
          }
          float res = (ar + as) + at;
          if (!bk) res = fmaf(g[GWJID * Np + n] * lambda[(int64_t)e * Np + n + offset], s_q[n], res);
          Aqe[n] = res;
        }
      }
    }
    free(s_q); free(s_r); free(s_s); free(s_t);
  }
}

/* axhelm/axhelmReference.cpp:19-101, plain += accumulation in its order */
void nbo_axhelm_reference(int Nq, int numElements, double lambda, const double* ggeo,
                          const double* D, const double* solIn, double* solOut)
{
  const int Nq2 = Nq * Nq, Np = Nq2 * Nq;
  double* Gr = (double*)malloc(sizeof(double) * Np);
  double* Gs = (double*)malloc(sizeof(double) * Np);
  double* Gt = (double*)malloc(sizeof(double) * Np);
  for (int e = 0; e < numElements; ++e) {
    const double* q = solIn + (int64_t)e * Np;
    double* out = solOut + (int64_t)e * Np;
    const double* g = ggeo + (int64_t)e * NGGEO * Np;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          double qr = 0, qs = 0, qt = 0;
          for (int n = 0; n < Nq; ++n) {
            qr += D[n + i * Nq] * q[n + j * Nq + k * Nq2];
            qs += D[n + j * Nq] * q[i + n * Nq + k * Nq2];
            qt += D[n + k * Nq] * q[i + j * Nq + n * Nq2];
          }
          const int id = i + j * Nq + k * Nq2;
          const double G00 = g[G00ID * Np + id], G01 = g[G01ID * Np + id], G02 = g[G02ID * Np + id];
          const double G11 = g[G11ID * Np + id], G12 = g[G12ID * Np + id], G22 = g[G22ID * Np + id];
          Gr[id] = (G00 * qr + G01 * qs + G02 * qt);
          Gs[id] = (G01 * qr + G11 * qs + G12 * qt);
          Gt[id] = (G02 * qr + G12 * qs + G22 * qt);
        }
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int id = i + j * Nq + k * Nq2;
          double lapq = lambda * g[GWJID * Np + id] * q[id];
          for (int n = 0; n < Nq; ++n) {
            lapq += D[i + n * Nq] * Gr[n + j * Nq + k * Nq2];
            lapq += D[j + n * Nq] * Gs[i + n * Nq + k * Nq2];
            lapq += D[k + n * Nq] * Gt[i + j * Nq + n * Nq2];
          }
          out[id] = lapq;
        }
  }
  free(Gr); free(Gs); free(Gt);
}

/* ------------------------------------------------------------------------ */
/* updateJacobi (nekBone/kernel/updateJacobi.okl:27-108), same expression
 * order; mapB[id]==1 marks a masked node whose diagonal is 1. */
void nbo_jacobi_local(int N, int32_t Nelements, int32_t offset, const int32_t* mapB,
                      const double* ggeo, const double* D, const double* lambda, double* diag)
{
  const int Nq = N + 1, Nq2 = Nq * Nq, Np = Nq2 * Nq;
#pragma omp parallel for schedule(static)
  for (int32_t e = 0; e < Nelements; ++e) {
    const double* g = ggeo + (int64_t)e * NGGEO * Np;
    const double* l0 = lambda + (int64_t)e * Np;
    for (int k = 0; k < Nq; ++k)
      for (int j = 0; j < Nq; ++j)
        for (int i = 0; i < Nq; ++i) {
          const int n = i + j * Nq + k * Nq2;
          const int64_t id = (int64_t)e * Np + n;
          double r_q = 1.0;
          if (!(mapB && mapB[id] == 1)) {
            r_q = 0.0;
            const double lambda_0 = l0[n];
            r_q += 2.0 * g[G01ID * Np + n] * lambda_0 * D[i * Nq + i] * D[j * Nq + j];
            r_q += 2.0 * g[G02ID * Np + n] * lambda_0 * D[i * Nq + i] * D[k * Nq + k];
            r_q += 2.0 * g[G12ID * Np + n] * lambda_0 * D[j * Nq + j] * D[k * Nq + k];
            for (int m = 0; m < Nq; ++m) {
              const int nr = m + j * Nq + k * Nq2, ns = i + m * Nq + k * Nq2, nt = i + j * Nq + m * Nq2;
              r_q += g[G00ID * Np + nr] * l0[nr] * D[m * Nq + i] * D[m * Nq + i];
              r_q += g[G11ID * Np + ns] * l0[ns] * D[m * Nq + j] * D[m * Nq + j];
              r_q += g[G22ID * Np + nt] * l0[nt] * D[m * Nq + k] * D[m * Nq + k];
            }
            r_q += g[GWJID * Np + n] * lambda[id + offset];
          }
          diag[id] = r_q;
        }
  }
}

/* ------------------------------------------------------------------------ */
/* The library's reduction tree (also what the HIP kernels do):
 *   tile of 2048 = 256 lanes x 8, lane t of tile b takes b*2048 + v*256 + t
 *   lane value: fma chain over v ascending
 *   tree over 256 lanes: inside each 64-lane wave a halving tree
 *   (a[t] += a[t+h], h = 32..1), then (S0 + S2) + (S1 + S3)
 *   stage 2: lane t adds partials t, t+256, ... sequentially, same tree. */
double nbo_tree256(const double* a)
{
  double S[4];
  double tmp[64];
  for (int w = 0; w < 4; ++w) {
    for (int t = 0; t < 64; ++t) tmp[t] = a[w * 64 + t];
    for (int h = 32; h >= 1; h >>= 1)
      for (int t = 0; t < h; ++t) tmp[t] = tmp[t] + tmp[t + h];
    S[w] = tmp[0];
  }
  return (S[0] + S[2]) + (S[1] + S[3]);
}

double nbo_tree_sum_partials(int64_t nb, const double* p)
{
  double lane[256];
  for (int t = 0; t < 256; ++t) {
    double acc = 0.0;
    for (int64_t i = t; i < nb; i += 256) acc = acc + p[i];
    lane[t] = acc;
  }
  return nbo_tree256(lane);
}

static double tree_dot_impl(int64_t n, const double* w, const double* x, const double* y)
{
  const int64_t nb = (n + 2047) / 2048;
  double* part = (double*)malloc(sizeof(double) * (nb > 0 ? nb : 1));
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) {
    double lane[256];
    for (int t = 0; t < 256; ++t) {
      double acc = 0.0;
      for (int v = 0; v < 8; ++v) {
        const int64_t i = b * 2048 + v * 256 + t;
        if (i < n) acc = w ? fma(w[i] * x[i], y[i], acc) : fma(x[i], y[i], acc);
      }
      lane[t] = acc;
    }
    part[b] = nbo_tree256(lane);
  }
  const double r = nbo_tree_sum_partials(nb, part);
  free(part);
  return r;
}

double nbo_tree_wdot(int64_t n, const double* w, const double* x, const double* y)
{
  return tree_dot_impl(n, w, x, y);
}

double nbo_tree_dot(int64_t n, const double* x, const double* y)
{
  return tree_dot_impl(n, NULL, x, y);
}

#include <omp.h>
int nbo_max_threads(void) { return omp_get_max_threads(); }

/* axhelm with matrix-core contractions (library: k_axhelm_mma and
 * k_axhelm_mma16, nbx_axhelm_mma.hip), any N, Nf fields: q, D and the G-scaled gradients are
 * rounded to the operand type (bf16 or fp16, nearest even); the 8-term
 * contractions are accumulated in double and rounded to float (the products
 * are exact in FP32; the matrix cores' internal summation order is not
 * specified, so agreement is to FP32 rounding, not bit for bit); geometric
 * factors applied in FP32 with the FP32 kernel's fma order; direction
 * results summed r, then s, then t. */
float nbo_bf16r(float x)
{
  union { float f; uint32_t u; } a;
  a.f = x;
  if ((a.u & 0x7FFFFFFFu) > 0x7F800000u) { /* NaN: keep sign and leading payload, set the quiet bit */
    a.u = (a.u | 0x00400000u) & 0xFFFF0000u;
    return a.f;
  }
  /* adding 0x7FFF (+1 when the kept part is odd) carries into bit 16 exactly
   * when the dropped half is above the midpoint, or at it with an odd kept part */
  a.u = (a.u + 0x7FFFu + ((a.u >> 16) & 1u)) & 0xFFFF0000u;
  return a.f;
}

void nbo_bf16_round_array(int64_t n, const float* x, float* y)
{
  for (int64_t i = 0; i < n; ++i) y[i] = nbo_bf16r(x[i]);
}

static float bf16r(float x) { return nbo_bf16r(x); }

/* IEEE binary16 rounding (nearest even) of a float, returned as a float:
 * 10 fraction bits, normal range from 2^-14, subnormal quantum 2^-24,
 * overflow to infinity from 65520 (half-way between 65504 and 2^16).  Every
 * scaling is by a power of two and nearbyintf rounds to nearest even, so
 * each step is exact. */
static float f16r(float x)
{
  const float ax = fabsf(x);
  if (!(ax < 65520.0f)) return isnan(x) ? x : copysignf(INFINITY, x);
  if (ax < 0x1p-14f) return nearbyintf(x * 0x1p24f) * 0x1p-24f;
  int ex;
  frexpf(x, &ex); /* |x| in [2^(ex-1), 2^ex): quantum 2^(ex-11) */
  return ldexpf(nearbyintf(ldexpf(x, 11 - ex)), ex - 11);
}

void nbo_f16_round_array(int64_t n, const float* x, float* y)
{
  for (int64_t i = 0; i < n; ++i) y[i] = f16r(x[i]);
}

static void axhelm_mma_impl(float (*h)(float), int N, int Nf, int bk, int32_t E, int32_t offset, const float* ggeo,
                            const float* D, const float* lambda, const float* q, float* Aq);

void nbo_axhelm_f16mma(int bk, int32_t E, int32_t offset, const float* ggeo, const float* D, const float* lambda,
                       const float* q, float* Aq)
{
  axhelm_mma_impl(f16r, 7, 1, bk, E, offset, ggeo, D, lambda, q, Aq);
}

void nbo_axhelm_bf16mma(int bk, int32_t E, int32_t offset, const float* ggeo, const float* D, const float* lambda,
                        const float* q, float* Aq)
{
  axhelm_mma_impl(bf16r, 7, 1, bk, E, offset, ggeo, D, lambda, q, Aq);
}

/* any order N <= 15 and Nf fields (q, Aq: field f at f*offset; the same
 * ggeo and lambda for every field, as the three-field kernels) */
void nbo_axhelm_mma_n(int f16, int N, int Nf, int bk, int32_t E, int32_t offset, const float* ggeo, const float* D,
                      const float* lambda, const float* q, float* Aq)
{
  axhelm_mma_impl(f16 ? f16r : bf16r, N, Nf, bk, E, offset, ggeo, D, lambda, q, Aq);
}

static void axhelm_mma_impl(float (*h)(float), int N, int Nf, int bk, int32_t E, int32_t offset, const float* ggeo,
                            const float* D, const float* lambda, const float* q, float* Aq)
{
  enum { QM = 16, PM = QM * QM * QM };
  const int Q = N + 1, P = Q * Q * Q;
  float Db[QM * QM];
  for (int i = 0; i < Q * Q; ++i) Db[i] = h(D[i]);
#pragma omp parallel
  {
    float qb[PM], qr[PM], qs[PM], qt[PM], gr[PM], gs[PM], gt[PM];
#pragma omp for schedule(static) collapse(2)
    for (int32_t e = 0; e < E; ++e)
    for (int f = 0; f < Nf; ++f) {
      const float* qe = q + (int64_t)f * offset + (int64_t)e * P;
      float* ae = Aq + (int64_t)f * offset + (int64_t)e * P;
      const float* g = ggeo + (int64_t)e * 7 * P;
      for (int n = 0; n < P; ++n) qb[n] = h(qe[n]);
      for (int n = 0; n < P; ++n) {
        const int i = n % Q, j = (n / Q) % Q, k = n / (Q * Q);
        double ar = 0, as = 0, at = 0;
        for (int m = 0; m < Q; ++m) {
          ar += (double)Db[i * Q + m] * qb[m + Q * j + Q * Q * k];
          as += (double)Db[j * Q + m] * qb[i + Q * m + Q * Q * k];
          at += (double)Db[k * Q + m] * qb[i + Q * j + Q * Q * m];
        }
        qr[n] = (float)ar; qs[n] = (float)as; qt[n] = (float)at;
      }
      for (int n = 0; n < P; ++n) {
        const float G00 = g[1 * P + n], G01 = g[2 * P + n], G02 = g[3 * P + n];
        const float G11 = g[4 * P + n], G12 = g[5 * P + n], G22 = g[6 * P + n];
        float Gr = fmaf(G02, qt[n], fmaf(G01, qs[n], G00 * qr[n]));
        float Gs = fmaf(G12, qt[n], fmaf(G11, qs[n], G01 * qr[n]));
        float Gt = fmaf(G22, qt[n], fmaf(G12, qs[n], G02 * qr[n]));
        if (!bk) { const float l0 = lambda[(int64_t)e * P + n]; Gr = l0 * Gr; Gs = l0 * Gs; Gt = l0 * Gt; }
        gr[n] = h(Gr); gs[n] = h(Gs); gt[n] = h(Gt);
      }
      for (int n = 0; n < P; ++n) {
        const int i = n % Q, j = (n / Q) % Q, k = n / (Q * Q);
        double ar = 0, as = 0, at = 0;
        for (int m = 0; m < Q; ++m) {
          ar += (double)Db[m * Q + i] * gr[m + Q * j + Q * Q * k];
          as += (double)Db[m * Q + j] * gs[i + Q * m + Q * Q * k];
          at += (double)Db[m * Q + k] * gt[i + Q * j + Q * Q * m];
        }
        float res = ((float)ar + (float)as) + (float)at;
        if (!bk) res = fmaf(g[0 * P + n] * lambda[(int64_t)e * P + offset + n], qe[n], res);
        ae[n] = res;
      }
    }
  }
}
