// Host-side producers of the hot path's inputs: GLL basis, box mesh, geometric
// factors and global node ids, as the reference's mesh setup builds them
// (core/meshSetup.cpp:2265-2451 box, :2156-2233 trilinear nodes, :266-360
// ggeo, :1638-1735 global ids; core/meshBasis.cpp:140-176, 775-790 basis).
//
// Differences from the reference, all deliberate and parameterised:
//   - the box may be bounded (Dirichlet) as well as periodic;
//   - interior vertices may be perturbed (seeded splitmix64 stream 1) and
//     z planes geometrically graded;
//   - elements stay in lexicographic order e = ex + NX*(ey + NY*ez) (the
//     reference sorts them by Morton index, core/meshSetup.cpp:531-779).
// Compiled with -ffp-contract=off: the expression order below is the
// reference's, so nodes and ggeo are reproducible bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "../../include/nbx.h"

namespace nbx {
int fail(int code, const char* fmt, ...);
}

namespace {

inline uint64_t splitmix64(uint64_t x)
{
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Urand {
  uint64_t key;
  Urand(uint64_t seed, uint64_t stream) : key(splitmix64(seed ^ splitmix64(stream))) {}
  double operator()(uint64_t i) const { return (double)(splitmix64(key + i) >> 11) * 0x1.0p-52 - 1.0; }
};

// P_N(x), P_N'(x)
void legendre(int N, double x, double& P, double& dP)
{
  if (N == 0) { P = 1.0; dP = 0.0; return; }
  double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
  for (int n = 2; n <= N; ++n) {
    const double p2 = ((2.0 * n - 1.0) * x * p1 - (n - 1.0) * p0) / n;
    const double d2 = d0 + (2.0 * n - 1.0) * p1;
    p0 = p1; p1 = p2; d0 = d1; d1 = d2;
  }
  P = p1; dP = d1;
}

void gll(int N, double* z, double* w)
{
  z[0] = -1.0;
  z[N] = 1.0;
  for (int i = 1; 2 * i < N + 1; ++i) {
    if (2 * i == N) { z[i] = 0.0; continue; }
    double x = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; ++it) {
      double P, dP;
      legendre(N, x, P, dP);
      const double d2P = (2.0 * x * dP - N * (N + 1.0) * P) / (1.0 - x * x);
      const double dx = dP / d2P;
      x -= dx;
      if (std::fabs(dx) < 1e-17) break;
    }
    z[i] = x;
    z[N - i] = -x;
  }
  for (int i = 0; i <= N; ++i) {
    double P, dP;
    legendre(N, z[i], P, dP);
    w[i] = 2.0 / (N * (N + 1.0) * P * P);
  }
}

void zplanes(const nbx_box& b, std::vector<double>& zp)
{
  const int NZ = b.NZ;
  zp.resize(NZ + 1);
  if (b.zaspect <= 0.0) {
    for (int k = 0; k <= NZ; ++k) zp[k] = b.z0 + (b.z1 - b.z0) * k / NZ;
    return;
  }
  const double Lz = b.z1 - b.z0, hx = (b.x1 - b.x0) / b.NX, dz0 = hx / b.zaspect;
  double lo = 1.0 + 1e-12, hi = 4.0;
  for (int it = 0; it < 200; ++it) {
    const double g = 0.5 * (lo + hi);
    if (dz0 * (std::pow(g, NZ) - 1.0) / (g - 1.0) > Lz) hi = g; else lo = g;
  }
  const double g = 0.5 * (lo + hi);
  for (int k = 0; k <= NZ; ++k) zp[k] = b.z0 + Lz * (std::pow(g, k) - 1.0) / (std::pow(g, NZ) - 1.0);
  zp[NZ] = b.z1;
}

}  // namespace

namespace nbx {

// Vertex coordinates of the box, 3 doubles per vertex (I, J, K) at
// 3 * (I + (NX+1) * (J + (NY+1) * K)): planes x0 + hx I, y0 + hy J, z-plane K;
// interior vertices moved by perturb * (hx, hy, hz) * U[-1,1) from stream 1,
// hz the smaller neighbouring z spacing.
void box_vertices(const nbx_box& b, std::vector<double>& V)
{
  std::vector<double> zp;
  zplanes(b, zp);
  const int NX = b.NX, NY = b.NY, NZ = b.NZ;
  const double hx = (b.x1 - b.x0) / NX, hy = (b.y1 - b.y0) / NY;
  const Urand vrand(b.seed, 1);
  V.resize((size_t)3 * (NX + 1) * (NY + 1) * (NZ + 1));
#pragma omp parallel for schedule(static)
  for (int K = 0; K <= NZ; ++K)
    for (int J = 0; J <= NY; ++J)
      for (int I = 0; I <= NX; ++I) {
        double px = b.x0 + hx * I, py = b.y0 + hy * J, pz = zp[K];
        const uint64_t vid = (uint64_t)(I + (int64_t)(NX + 1) * (J + (int64_t)(NY + 1) * K));
        if (b.perturb != 0.0 && I > 0 && I < NX && J > 0 && J < NY && K > 0 && K < NZ) {
          const double hz = std::min(zp[K] - zp[K - 1], zp[K + 1] - zp[K]);
          px += b.perturb * hx * vrand(3 * vid + 0);
          py += b.perturb * hy * vrand(3 * vid + 1);
          pz += b.perturb * hz * vrand(3 * vid + 2);
        }
        V[3 * vid + 0] = px;
        V[3 * vid + 1] = py;
        V[3 * vid + 2] = pz;
      }
}

// key of a splitmix64 uniform stream (the device fill uses it)
uint64_t urand_key(uint64_t seed, uint64_t stream) { return Urand(seed, stream).key; }

}  // namespace nbx

extern "C" int nbx_gll(int N, double* z, double* w, double* D)
{
  if (N < 1 || N > 31 || !z || !w) return nbx::fail(NBX_ERR_INVALID, "gll: bad arguments");
  gll(N, z, w);
  if (D) {
    const int Nq = N + 1;
    std::vector<double> P(Nq);
    for (int i = 0; i < Nq; ++i) { double dP; legendre(N, z[i], P[i], dP); }
    for (int i = 0; i < Nq; ++i)
      for (int j = 0; j < Nq; ++j) {
        double d = 0.0;
        if (i != j) d = P[i] / (P[j] * (z[i] - z[j]));
        else if (i == 0) d = -N * (N + 1.0) / 4.0;
        else if (i == N) d = N * (N + 1.0) / 4.0;
        D[i * Nq + j] = d;
      }
  }
  return NBX_SUCCESS;
}

extern "C" int nbx_urand_fill(uint64_t seed, uint64_t stream, int64_t n, double* out)
{
  if (n < 0 || (n > 0 && !out)) return nbx::fail(NBX_ERR_INVALID, "urand_fill: bad arguments");
  const Urand u(seed, stream);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) out[i] = u((uint64_t)i);
  return NBX_SUCCESS;
}

extern "C" int nbx_box_mesh(const nbx_box* box, int mask_boundary, double* x, double* y, double* z,
                            double* ggeo, nbx_hlong* globalIds, double* minJout)
{
  if (!box) return nbx::fail(NBX_ERR_INVALID, "box_mesh: null box");
  const nbx_box b = *box;
  if (b.N < 1 || b.N > 15 || b.NX < 1 || b.NY < 1 || b.NZ < 1)
    return nbx::fail(NBX_ERR_INVALID, "box_mesh: bad sizes");
  if (b.periodic && (b.NX < 3 || b.NY < 3 || b.NZ < 3))
    return nbx::fail(NBX_ERR_INVALID, "box_mesh: periodic box needs >= 3 elements per direction");
  const int N = b.N, Nq = N + 1, Nq2 = Nq * Nq, Np = Nq2 * Nq;
  const int NX = b.NX, NY = b.NY, NZ = b.NZ;
  const int64_t E = (int64_t)NX * NY * NZ;
  if (E * Np > INT32_MAX) return nbx::fail(NBX_ERR_INVALID, "box_mesh: E*Np exceeds dlong range");
  std::vector<double> zr(Nq), wr(Nq), D(Nq2);
  nbx_gll(N, zr.data(), wr.data(), D.data());
  std::vector<double> V;
  nbx::box_vertices(b, V);
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  double minJ = 1e300;
#pragma omp parallel reduction(min : minJ)
  {
    std::vector<double> xe(Np), ye(Np), ze(Np);
#pragma omp for schedule(static)
    for (int64_t e = 0; e < E; ++e) {
      const int ex = (int)(e % NX), ey = (int)((e / NX) % NY), ez = (int)(e / ((int64_t)NX * NY));
      double vx[8], vy[8], vz[8];
      for (int v = 0; v < 8; ++v) {
        const int I = ex + off[v][0], J = ey + off[v][1], K = ez + off[v][2];
        const size_t vid = (size_t)I + (size_t)(NX + 1) * (J + (size_t)(NY + 1) * K);
        vx[v] = V[3 * vid + 0]; vy[v] = V[3 * vid + 1]; vz[v] = V[3 * vid + 2];
      }
      // trilinear map, core/meshSetup.cpp:2198-2230 expression order
      for (int n = 0; n < Np; ++n) {
        const double rn = zr[n % Nq], sn = zr[(n / Nq) % Nq], tn = zr[n / Nq2];
        const double c0 = 0.125 * (1 - rn) * (1 - sn) * (1 - tn), c1 = 0.125 * (1 + rn) * (1 - sn) * (1 - tn);
        const double c2 = 0.125 * (1 + rn) * (1 + sn) * (1 - tn), c3 = 0.125 * (1 - rn) * (1 + sn) * (1 - tn);
        const double c4 = 0.125 * (1 - rn) * (1 - sn) * (1 + tn), c5 = 0.125 * (1 + rn) * (1 - sn) * (1 + tn);
        const double c6 = 0.125 * (1 + rn) * (1 + sn) * (1 + tn), c7 = 0.125 * (1 - rn) * (1 + sn) * (1 + tn);
        xe[n] = +c0 * vx[0] + c1 * vx[1] + c2 * vx[2] + c3 * vx[3] + c4 * vx[4] + c5 * vx[5] + c6 * vx[6] + c7 * vx[7];
        ye[n] = +c0 * vy[0] + c1 * vy[1] + c2 * vy[2] + c3 * vy[3] + c4 * vy[4] + c5 * vy[5] + c6 * vy[6] + c7 * vy[7];
        ze[n] = +c0 * vz[0] + c1 * vz[1] + c2 * vz[2] + c3 * vz[3] + c4 * vz[4] + c5 * vz[5] + c6 * vz[6] + c7 * vz[7];
      }
      if (x) std::memcpy(x + e * Np, xe.data(), sizeof(double) * Np);
      if (y) std::memcpy(y + e * Np, ye.data(), sizeof(double) * Np);
      if (z) std::memcpy(z + e * Np, ze.data(), sizeof(double) * Np);
      if (!ggeo) continue;
      // geometric factors, core/meshSetup.cpp:266-360
      double* g = ggeo + e * (int64_t)7 * Np;
      for (int k = 0; k < Nq; ++k)
        for (int j = 0; j < Nq; ++j)
          for (int i = 0; i < Nq; ++i) {
            const int n = i + j * Nq + k * Nq2;
            double xr = 0, xs = 0, xt = 0, yr = 0, ys = 0, yt = 0, zr_ = 0, zs = 0, zt = 0;
            for (int m = 0; m < Nq; ++m) {
              const int idr = k * Nq2 + j * Nq + m, ids = k * Nq2 + m * Nq + i, idt = m * Nq2 + j * Nq + i;
              const double Di = D[i * Nq + m], Dj = D[j * Nq + m], Dk = D[k * Nq + m];
              xr += Di * xe[idr]; xs += Dj * xe[ids]; xt += Dk * xe[idt];
              yr += Di * ye[idr]; ys += Dj * ye[ids]; yt += Dk * ye[idt];
              zr_ += Di * ze[idr]; zs += Dj * ze[ids]; zt += Dk * ze[idt];
            }
            const double J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr_ * (xs * yt - ys * xt);
            minJ = std::min(minJ, J);
            const double rx = (ys * zt - zs * yt) / J, ry = -(xs * zt - zs * xt) / J, rz = (xs * yt - ys * xt) / J;
            const double sx = -(yr * zt - zr_ * yt) / J, sy = (xr * zt - zr_ * xt) / J, sz = -(xr * yt - yr * xt) / J;
            const double tx = (yr * zs - zr_ * ys) / J, ty = -(xr * zs - zr_ * xs) / J, tz = (xr * ys - yr * xs) / J;
            const double JW = J * wr[i] * wr[j] * wr[k];
            g[0 * Np + n] = JW;
            g[1 * Np + n] = JW * (rx * rx + ry * ry + rz * rz);
            g[2 * Np + n] = JW * (rx * sx + ry * sy + rz * sz);
            g[3 * Np + n] = JW * (rx * tx + ry * ty + rz * tz);
            g[4 * Np + n] = JW * (sx * sx + sy * sy + sz * sz);
            g[5 * Np + n] = JW * (sx * tx + sy * ty + sz * tz);
            g[6 * Np + n] = JW * (tx * tx + ty * ty + tz * tz);
          }
    }
  }
  if (minJout) *minJout = minJ;
  if (!globalIds) return NBX_SUCCESS;
  // global ids: fixed point of meshParallelConnectNodes' min-label propagation
  // (vertex nodes: vertex number + 1; others: 1 + Nnodes + smallest local index
  // of the group).  Groups = equal structured GLL index.
  const int64_t GX = b.periodic ? (int64_t)NX * N : (int64_t)NX * N + 1;
  const int64_t GY = b.periodic ? (int64_t)NY * N : (int64_t)NY * N + 1;
  const int64_t GZ = b.periodic ? (int64_t)NZ * N : (int64_t)NZ * N + 1;
  const int64_t Nnodes = b.periodic ? E : (int64_t)(NX + 1) * (NY + 1) * (NZ + 1);
  std::vector<int64_t> first(GX * GY * GZ, -1);
  for (int64_t e = 0; e < E; ++e) {
    const int ex = (int)(e % NX), ey = (int)((e / NX) % NY), ez = (int)(e / ((int64_t)NX * NY));
    for (int n = 0; n < Np; ++n) {
      const int i = n % Nq, j = (n / Nq) % Nq, k = n / Nq2;
      int64_t I = (int64_t)ex * N + i, J = (int64_t)ey * N + j, K = (int64_t)ez * N + k;
      const bool onbnd = I == 0 || J == 0 || K == 0 || I == (int64_t)NX * N || J == (int64_t)NY * N ||
                         K == (int64_t)NZ * N;
      if (b.periodic) { I %= GX; J %= GY; K %= GZ; }
      const int64_t gidx = I + GX * (J + GY * K);
      const int64_t id = e * Np + n;
      if (first[gidx] < 0) first[gidx] = id;
      int64_t label;
      if ((i == 0 || i == N) && (j == 0 || j == N) && (k == 0 || k == N)) {
        const int64_t vi = ex + (i == N), vj = ey + (j == N), vk = ez + (k == N);
        label = 1 + (b.periodic ? (vi % NX) + (int64_t)NX * ((vj % NY) + (int64_t)NY * (vk % NZ))
                                : vi + (int64_t)(NX + 1) * (vj + (int64_t)(NY + 1) * vk));
      } else {
        label = 1 + first[gidx] + Nnodes;
      }
      globalIds[id] = (mask_boundary && !b.periodic && onbnd) ? 0 : label;
    }
  }
  return NBX_SUCCESS;
}
