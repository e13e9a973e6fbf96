// Burgers benchmark (Parthenon-VIBE) stage kernels for gfx950, FP64.
//
// Task-level kernels, one launch per reference task over every block of the MeshData:
//   ReconKernel      burgers_package::CalculateFluxes, reconstruction part
//                    (benchmarks/burgers/burgers_package.cpp:236-303, recon.hpp:28-99)
//   FluxKernel       the HLL part (burgers_package.cpp:307-401, burgers_package.hpp:33-43)
//   FluxDivKernel    Update::FluxDivergence (src/interface/update.cpp:63-84, update.hpp:43-57)
//   WsumKernel       Update::WeightedSumData (src/interface/update.hpp:71-90)
//   DerivedKernel    burgers_package::CalculateDerived (burgers_package.cpp:143-167)
//   DtKernel         burgers_package::EstimateTimestepMesh (burgers_package.cpp:170-200):
//                    per-workgroup min via 64-lane wavefront shuffles, then one final pass
// Expression order matches the reference operator by operator and the library is built with
// -ffp-contract=off, so results are reproducible bit for bit by an IEEE CPU evaluation.
// Arrays are [block][comp][k][j][i] bricks of (n+2g)^3 with i fastest; consecutive lanes
// take consecutive i, so every load and store is a coalesced run along x.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>

#include <algorithm>
#include <cfloat>

#include "pth_burgers.hpp"
#include "pth_device.hpp"

namespace pth {

namespace {

constexpr int kWG = 256;

struct Geo {   // brick shape
  int e0, e1, e2;       // entire extents
  int lo0, lo1, lo2;    // interior starts
  int hi0, hi1, hi2;    // interior ends
  int V;                // components
  int ndim;
  std::int64_t cells;   // e0*e1*e2
};

__device__ __forceinline__ double DMin(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double DMax(double a, double b) { return a < b ? b : a; }

__device__ __forceinline__ double Mc(double dm, double dp) {
  const double dc = (dm * dp > 0.0) * 0.5 * (dm + dp);
  return copysign(DMin(fabs(dc), 2.0 * DMin(fabs(dm), fabs(dp))), dc);
}

__device__ __forceinline__ void LinearRecon(double qm, double q0, double qp, double &ql,
                                            double &qr) {
  double dq = qp - q0;
  dq = 0.5 * Mc(q0 - qm, dq);
  ql = q0 + dq;
  qr = q0 - dq;
}

__device__ __forceinline__ void Weno5Z(double q0, double q1, double q2, double q3, double q4,
                                       double &ql, double &qr) {
  const double a00 = 1.0 / 3.0, a01 = -7.0 / 6.0, a02 = 11.0 / 6.0;
  const double a10 = -1.0 / 6.0, a11 = 5.0 / 6.0, a12 = 1.0 / 3.0;
  const double a20 = 1.0 / 3.0, a21 = 5.0 / 6.0, a22 = -1.0 / 6.0;
  const double g0 = 0.1, g1 = 0.6, g2 = 0.3;
  const double eps = 10.0 * DBL_EPSILON;  // robust::EPS (src/utils/robust.hpp:38-41)
  const double t13 = 13.0 / 3.0;
  double a = q0 - 2 * q1 + q2;
  double b = q0 - 4.0 * q1 + 3.0 * q2;
  double be0 = t13 * a * a + b * b + eps;
  a = q1 - 2.0 * q2 + q3;
  b = q3 - q1;
  double be1 = t13 * a * a + b * b + eps;
  a = q2 - 2.0 * q3 + q4;
  b = q4 - 4.0 * q3 + 3.0 * q2;
  double be2 = t13 * a * a + b * b + eps;
  const double tau5 = fabs(be2 - be0);
  be0 = (be0 + tau5) / be0;
  be1 = (be1 + tau5) / be1;
  be2 = (be2 + tau5) / be2;
  double w0 = g0 * be0 + eps;
  double w1 = g1 * be1 + eps;
  double w2 = g2 * be2 + eps;
  double ws = 1.0 / (w0 + w1 + w2);
  ql = w0 * (a00 * q0 + a01 * q1 + a02 * q2);
  ql += w1 * (a10 * q1 + a11 * q2 + a12 * q3);
  ql += w2 * (a20 * q2 + a21 * q3 + a22 * q4);
  ql *= ws;
  const double al = 3.0 * ws * w0 * w1 * w2 / (g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;
  w0 = g0 * be2 + eps;
  w1 = g1 * be1 + eps;
  w2 = g2 * be0 + eps;
  ws = 1.0 / (w0 + w1 + w2);
  qr = w0 * (a00 * q4 + a01 * q3 + a02 * q2);
  qr += w1 * (a10 * q3 + a11 * q2 + a12 * q1);
  qr += w2 * (a20 * q2 + a21 * q1 + a22 * q0);
  qr *= ws;
  const double ar = 3.0 * ws * w0 * w1 * w2 / (g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;
  double dq = q3 - q2;
  dq = 0.5 * Mc(q2 - q1, dq);
  const double alin = 2.0 * al * ar / (al + ar);
  ql = alin * ql + (1.0 - alin) * (q2 + dq);
  qr = alin * qr + (1.0 - alin) * (q2 - dq);
}

// One thread per (block, comp, k, j, i) of the box extended by one cell around the interior.
// face layout: [6][block][comp][k][j][i] for Ulx, Urx, Uly, Ury, Ulz, Urz.
__global__ void __launch_bounds__(kWG)
    ReconKernel(const double *__restrict__ U, double *__restrict__ face, Geo G, int nblocks,
                int recon) {
  const int dk = G.ndim > 2 ? 1 : 0, dj = G.ndim > 1 ? 1 : 0;
  const int b0 = G.hi0 - G.lo0 + 3, b1 = G.hi1 - G.lo1 + 1 + 2 * dj,
            b2 = G.hi2 - G.lo2 + 1 + 2 * dk;
  const std::int64_t per = static_cast<std::int64_t>(b0) * b1 * b2;
  const std::int64_t total = per * G.V * nblocks;
  const std::int64_t slab = G.cells * G.V * nblocks;
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<std::int64_t>(gridDim.x) * kWG) {
    std::int64_t r = t;
    const int i = static_cast<int>(r % b0) + G.lo0 - 1;
    r /= b0;
    const int j = static_cast<int>(r % b1) + G.lo1 - dj;
    r /= b1;
    const int k = static_cast<int>(r % b2) + G.lo2 - dk;
    r /= b2;  // r = block * V + comp
    const double *q = U + r * G.cells;
    const std::int64_t c = (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
    double *f = face + r * G.cells;
    const bool kin = k >= G.lo2 && k <= G.hi2, jin = j >= G.lo1 && j <= G.hi1,
               iin = i >= G.lo0 && i <= G.hi0;
    const bool xrec = kin && jin;
    const bool yrec = kin && (G.ndim > 1) && iin;
    const bool zrec = jin && (G.ndim > 2) && iin;
    const std::int64_t sj = G.e0, sk = static_cast<std::int64_t>(G.e0) * G.e1;
    double ql, qr;
    if (xrec) {
      if (recon == 0)
        Weno5Z(q[c - 2], q[c - 1], q[c], q[c + 1], q[c + 2], ql, qr);
      else
        LinearRecon(q[c - 1], q[c], q[c + 1], ql, qr);
      f[0 * slab + c + 1] = ql;
      f[1 * slab + c] = qr;
    }
    if (yrec) {
      if (recon == 0)
        Weno5Z(q[c - 2 * sj], q[c - sj], q[c], q[c + sj], q[c + 2 * sj], ql, qr);
      else
        LinearRecon(q[c - sj], q[c], q[c + sj], ql, qr);
      f[2 * slab + c + sj] = ql;
      f[3 * slab + c] = qr;
    }
    if (zrec) {
      if (recon == 0)
        Weno5Z(q[c - 2 * sk], q[c - sk], q[c], q[c + sk], q[c + 2 * sk], ql, qr);
      else
        LinearRecon(q[c - sk], q[c], q[c + sk], ql, qr);
      f[4 * slab + c + sk] = ql;
      f[5 * slab + c] = qr;
    }
  }
}

// One thread per (block, k, j, i) face position; HLL velocity flux then scalar fluxes that
// reuse the face's signal speeds. flux layout: [3][block][comp][k][j][i].
__global__ void __launch_bounds__(kWG)
    FluxKernel(const double *__restrict__ face, double *__restrict__ flux, Geo G, int nblocks) {
  const int dk = G.ndim > 2 ? 1 : 0, dj = G.ndim > 1 ? 1 : 0;
  const int b0 = G.hi0 - G.lo0 + 2, b1 = G.hi1 - G.lo1 + 1 + dj, b2 = G.hi2 - G.lo2 + 1 + dk;
  const std::int64_t per = static_cast<std::int64_t>(b0) * b1 * b2;
  const std::int64_t total = per * nblocks;
  const std::int64_t slab = G.cells * G.V * nblocks;
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<std::int64_t>(gridDim.x) * kWG) {
    std::int64_t r = t;
    const int i = static_cast<int>(r % b0) + G.lo0;
    r /= b0;
    const int j = static_cast<int>(r % b1) + G.lo1;
    r /= b1;
    const int k = static_cast<int>(r % b2) + G.lo2;
    r /= b2;  // block
    const std::int64_t c = (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
    const std::int64_t base = r * G.V * G.cells + c;
    const bool xflux = k <= G.hi2 && j <= G.hi1;
    const bool yflux = G.ndim > 1 && k <= G.hi2 && i <= G.hi0;
    const bool zflux = G.ndim > 2 && j <= G.hi1 && i <= G.hi0;
    for (int d = 0; d < 3; ++d) {
      const bool doit = d == 0 ? xflux : (d == 1 ? yflux : zflux);
      if (!doit) continue;
      const double *L = face + (2 * d) * slab + base;
      const double *R = face + (2 * d + 1) * slab + base;
      double *F = flux + d * slab + base;
      const double upl = L[d * G.cells], upr = R[d * G.cells];
      const double sl = DMin(DMin(upl, upr), 0.0);
      const double sr = DMax(DMax(upl, upr), 0.0);
      const double islsr = 1.0 / (sr - sl + (sl * sr == 0.0));
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        const double ul = L[n * G.cells], ur = R[n * G.cells];
        F[n * G.cells] = 0.5 * (sr * ul * upl - sl * ur * upr + sl * sr * (ur - ul)) * islsr;
      }
      for (int n = 3; n < G.V; ++n) {
        const double ql = L[n * G.cells], qr = R[n * G.cells];
        F[n * G.cells] =
            (sr * upl * ql - sl * upr * qr + sl * sr * (qr - ql)) / (sr - sl + (sl * sr == 0.0));
      }
    }
  }
}

__global__ void __launch_bounds__(kWG)
    FluxDivKernel(const double *__restrict__ flux, double *__restrict__ dudt,
                  const BlockCoefs *__restrict__ coef, Geo G, int nblocks) {
  const int n0 = G.hi0 - G.lo0 + 1, n1 = G.hi1 - G.lo1 + 1, n2 = G.hi2 - G.lo2 + 1;
  const std::int64_t total = static_cast<std::int64_t>(n0) * n1 * n2 * G.V * nblocks;
  const std::int64_t slab = G.cells * G.V * nblocks;
  const std::int64_t sj = G.e0, sk = static_cast<std::int64_t>(G.e0) * G.e1;
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<std::int64_t>(gridDim.x) * kWG) {
    std::int64_t r = t;
    const int i = static_cast<int>(r % n0) + G.lo0;
    r /= n0;
    const int j = static_cast<int>(r % n1) + G.lo1;
    r /= n1;
    const int k = static_cast<int>(r % n2) + G.lo2;
    r /= n2;  // block * V + comp
    const int b = static_cast<int>(r / G.V);
    const BlockCoefs &C = coef[b];
    const std::int64_t c = r * G.cells + (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
    const double *fx = flux + c, *fy = flux + slab + c, *fz = flux + 2 * slab + c;
    double du = (C.area[0] * fx[1] - C.area[0] * fx[0]);
    if (G.ndim >= 2) du += (C.area[1] * fy[sj] - C.area[1] * fy[0]);
    if (G.ndim == 3) du += (C.area[2] * fz[sk] - C.area[2] * fz[0]);
    dudt[c] = -du / C.vol;
  }
}

__global__ void __launch_bounds__(kWG)
    WsumKernel(const double *__restrict__ x, const double *__restrict__ y, double w1, double w2,
               double *__restrict__ z, std::int64_t n) {
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < n;
       t += static_cast<std::int64_t>(gridDim.x) * kWG)
    z[t] = w1 * x[t] + w2 * y[t];
}

__global__ void __launch_bounds__(kWG)
    DerivedKernel(const double *__restrict__ U, double *__restrict__ out, Geo G, int nblocks) {
  const int n0 = G.hi0 - G.lo0 + 1, n1 = G.hi1 - G.lo1 + 1, n2 = G.hi2 - G.lo2 + 1;
  const std::int64_t total = static_cast<std::int64_t>(n0) * n1 * n2 * nblocks;
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<std::int64_t>(gridDim.x) * kWG) {
    std::int64_t r = t;
    const int i = static_cast<int>(r % n0) + G.lo0;
    r /= n0;
    const int j = static_cast<int>(r % n1) + G.lo1;
    r /= n1;
    const int k = static_cast<int>(r % n2) + G.lo2;
    r /= n2;
    const std::int64_t c = (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
    const double *u = U + r * G.V * G.cells + c;
    const double u1 = u[0], u2 = u[G.cells], u3 = u[2 * G.cells], d = u[3 * G.cells];
    out[r * G.cells + c] = 0.5 * d * (u1 * u1 + u2 * u2 + u3 * u3);
  }
}

__device__ __forceinline__ double WaveMin(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = DMin(v, __shfl_xor(v, off, 64));
  return v;
}

// Per-workgroup minimum of the CFL time step over interior cells; partial[blockIdx.x].
__global__ void __launch_bounds__(kWG)
    DtKernel(const double *__restrict__ U, const BlockCoefs *__restrict__ coef, Geo G,
             int nblocks, double *__restrict__ partial) {
  __shared__ double wmin[kWG / 64];
  const int n0 = G.hi0 - G.lo0 + 1, n1 = G.hi1 - G.lo1 + 1, n2 = G.hi2 - G.lo2 + 1;
  const std::int64_t total = static_cast<std::int64_t>(n0) * n1 * n2 * nblocks;
  double m = DBL_MAX;
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<std::int64_t>(gridDim.x) * kWG) {
    std::int64_t r = t;
    const int i = static_cast<int>(r % n0) + G.lo0;
    r /= n0;
    const int j = static_cast<int>(r % n1) + G.lo1;
    r /= n1;
    const int k = static_cast<int>(r % n2) + G.lo2;
    r /= n2;
    const BlockCoefs &C = coef[r];
    const double *u = U + r * G.V * G.cells + (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
    double s = fabs(u[0]) / C.dx[0];
    if (G.ndim > 1) s = s + fabs(u[G.cells]) / C.dx[1];
    if (G.ndim > 2) s = s + fabs(u[2 * G.cells]) / C.dx[2];
    m = DMin(m, 1.0 / s);
  }
  m = WaveMin(m);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = wmin[0];
#pragma unroll
    for (int w = 1; w < kWG / 64; ++w) v = DMin(v, wmin[w]);
    partial[blockIdx.x] = v;
  }
}

__global__ void __launch_bounds__(64) MinFinalKernel(const double *__restrict__ partial, int n,
                                                     double *__restrict__ out) {
  double m = DBL_MAX;
  for (int t = threadIdx.x; t < n; t += 64) m = DMin(m, partial[t]);
  m = WaveMin(m);
  if (threadIdx.x == 0) out[0] = m;
}

// ---- fused stage path ----------------------------------------------------------------
//
// ReconTaskKernel: one thread = one WENO5Z (or linear) reconstruction of one cell in one
// direction over one brick (block, comp). Per direction the task box
// is the interior extended by one cell along that direction only, so every lane does the same
// amount of work. The index work around the FP64
// reconstruction is kept small: the brick shape is a compile-time constant for the common
// cubic blocks (N interior cells, NG ghosts; N == 0 reads the shape at run time), the direction
// is uniform per workgroup, and loads and stores use 32-bit byte offsets from per-workgroup
// base pointers.
__device__ __forceinline__ double LoadAt(const double *base, unsigned byte_off) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + byte_off);
}
__device__ __forceinline__ void StoreAt(double *base, unsigned byte_off, double v) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

template <int N, int NG, int D>
__device__ __forceinline__ void ReconTask(const double *__restrict__ q, double *__restrict__ fl,
                                          double *__restrict__ fr, const Geo &G, int t,
                                          int recon) {
  const int n0 = N ? N : G.hi0 - G.lo0 + 1, n1 = N ? N : G.hi1 - G.lo1 + 1,
            n2 = N ? N : G.hi2 - G.lo2 + 1;
  const int lo0 = N ? NG : G.lo0, lo1 = N ? NG : G.lo1, lo2 = N ? NG : G.lo2;
  const int e0 = N ? N + 2 * NG : G.e0, e1 = N ? N + 2 * NG : G.e1;
  const int m0 = n0 + (D == 0 ? 2 : 0), m1 = n1 + (D == 1 ? 2 : 0), m2 = n2 + (D == 2 ? 2 : 0);
  if (t >= m0 * m1 * m2) return;
  const int i = t % m0 + lo0 - (D == 0 ? 1 : 0);
  const int r = t / m0;
  const int j = r % m1 + lo1 - (D == 1 ? 1 : 0);
  const int k = r / m1 + lo2 - (D == 2 ? 1 : 0);
  const unsigned c = 8u * static_cast<unsigned>((k * e1 + j) * e0 + i);
  const unsigned st = 8u * static_cast<unsigned>(D == 0 ? 1 : D == 1 ? e0 : e0 * e1);
  double ql, qr;
  if (recon == 0)
    Weno5Z(LoadAt(q, c - 2 * st), LoadAt(q, c - st), LoadAt(q, c), LoadAt(q, c + st),
           LoadAt(q, c + 2 * st), ql, qr);
  else
    LinearRecon(LoadAt(q, c - st), LoadAt(q, c), LoadAt(q, c + st), ql, qr);
  StoreAt(fl, c + st, ql);
  StoreAt(fr, c, qr);
}

template <int N, int NG>
__global__ void __launch_bounds__(kWG)
    ReconTaskKernel(const double *__restrict__ U, double *__restrict__ face, Geo G, int nblocks,
                    int recon, unsigned per_brick) {
  // logical workgroup = (block * V + comp) * per_brick + tile * ndim + direction; consecutive
  // logical workgroups take the directions of one tile and run on the same XCD, so the three
  // passes over the same cells of U meet in one L2
  const unsigned L = XcdRemap(blockIdx.x, gridDim.x);
  const std::int64_t bc = L / per_brick;
  const unsigned r = L % per_brick;
  const std::int64_t slab = G.cells * G.V * nblocks;
  const double *q = U + bc * G.cells;
  const int d = r % G.ndim;
  double *fl = face + 2 * d * slab + bc * G.cells;
  double *fr = fl + slab;
  const int t = (r / G.ndim) * kWG + threadIdx.x;
  if (d == 0)
    ReconTask<N, NG, 0>(q, fl, fr, G, t, recon);
  else if (d == 1)
    ReconTask<N, NG, 1>(q, fl, fr, G, t, recon);
  else
    ReconTask<N, NG, 2>(q, fl, fr, G, t, recon);
}

// HLL flux of every component at one face from the face states (burgers_package.cpp:307-401
// with lr_to_flux, burgers_package.hpp:33-43). l(n), r(n) give component n of the left and
// right states; d = normal direction.
template <int V, class LS, class RS>
__device__ __forceinline__ void HllGen(LS l, RS r, int d, double *F) {
  const double upl = l(d), upr = r(d);
  const double sl = DMin(DMin(upl, upr), 0.0);
  const double sr = DMax(DMax(upl, upr), 0.0);
  const double islsr = 1.0 / (sr - sl + (sl * sr == 0.0));
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    const double ul = l(n), ur = r(n);
    F[n] = 0.5 * (sr * ul * upl - sl * ur * upr + sl * sr * (ur - ul)) * islsr;
  }
#pragma unroll
  for (int n = 3; n < V; ++n) {
    const double ql = l(n), qr = r(n);
    F[n] = (sr * upl * ql - sl * upr * qr + sl * sr * (qr - ql)) / (sr - sl + (sl * sr == 0.0));
  }
}

// HllGen over state arrays: L, R point at component 0 of the face; component stride cs.
template <int V>
__device__ __forceinline__ void HllFace(const double *__restrict__ L, const double *__restrict__ R,
                                        std::int64_t cs, int d, double *F) {
  HllGen<V>([&](int n) { return L[n * cs]; }, [&](int n) { return R[n * cs]; }, d, F);
}

constexpr int kMaxV = 16;  // components handled by the fused update kernel

// FusedUpdateKernel: one thread = one interior cell. HLL fluxes on both faces of the cell in
// every direction, flux divergence (update.hpp:43-57), the RK averaging and update in the
// WeightedSumData form (update.hpp:71-90, 122-137), the derived field and the per-workgroup
// CFL minimum of the updated state; optionally writes the low-face fluxes (and the last high
// face) for flux correction. Reads face states written by ReconTaskKernel. On multilevel
// meshes `fover` gives per block face the correction-store image whose values replace the HLL
// fluxes of that face (SetBounds<flxcor_recv> before FluxDivergence in the reference order).
template <int V>
__global__ void __launch_bounds__(kWG)
    FusedUpdateKernel(const double *__restrict__ face, const double *U0, const double *base,
                      double *out, double *__restrict__ derived, double *__restrict__ flux,
                      double *__restrict__ partial, const BlockCoefs *__restrict__ coef, Geo G,
                      int nblocks, double beta, const double *__restrict__ dtvar,
                      const int *__restrict__ fover, const double *__restrict__ fcor) {
  __shared__ double wmin[kWG / 64];
  const int n0 = G.hi0 - G.lo0 + 1, n1 = G.hi1 - G.lo1 + 1, n2 = G.hi2 - G.lo2 + 1;
  const unsigned L = XcdRemap(blockIdx.x, gridDim.x);  // = block * tiles + tile
  const int tiles = (n0 * n1 * n2 + kWG - 1) / kWG;
  const int b = static_cast<int>(L / tiles);
  const int t = static_cast<int>(L % tiles) * kWG + threadIdx.x;
  double m = DBL_MAX;
  if (t < n0 * n1 * n2) {
    const int i = t % n0 + G.lo0;
    const int r = t / n0;
    const int j = r % n1 + G.lo1;
    const int k = r / n1 + G.lo2;
    const std::int64_t slab = G.cells * V * nblocks;
    const std::int64_t cs = G.cells;
    const std::int64_t c0 = static_cast<std::int64_t>(b) * V * cs + (k * G.e1 + j) * G.e0 + i;
    const BlockCoefs &C = coef[b];
    const std::int64_t st[3] = {1, G.e0, static_cast<std::int64_t>(G.e0) * G.e1};
    double du[V];
    for (int d = 0; d < G.ndim; ++d) {
      double Flo[V], Fhi[V];
      const double *L = face + (2 * d) * slab + c0, *R = face + (2 * d + 1) * slab + c0;
      HllFace<V>(L, R, cs, d, Flo);
      HllFace<V>(L + st[d], R + st[d], cs, d, Fhi);
      if (fover) {  // flux correction: a finer neighbour's restricted fluxes replace ours
        const int at = d == 0 ? i - G.lo0 : d == 1 ? j - G.lo1 : k - G.lo2;
        const int top = d == 0 ? G.hi0 - i : d == 1 ? G.hi1 - j : G.hi2 - k;
        if (at == 0 || top == 0) {
          const int a = d == 0 ? j - G.lo1 : i - G.lo0;
          const int q = d == 2 ? j - G.lo1 : k - G.lo2;
          const int rows = d == 0 ? n1 : n0;
          const int comp = rows * (d == 2 ? n1 : n2);
          const int olo = at == 0 ? fover[6 * b + 2 * d] : -1;
          const int ohi = top == 0 ? fover[6 * b + 2 * d + 1] : -1;
          if (olo >= 0)
#pragma unroll
            for (int n = 0; n < V; ++n) Flo[n] = fcor[olo + a + q * rows + n * comp];
          if (ohi >= 0)
#pragma unroll
            for (int n = 0; n < V; ++n) Fhi[n] = fcor[ohi + a + q * rows + n * comp];
        }
      }
      const double A = C.area[d];
#pragma unroll
      for (int n = 0; n < V; ++n) {
        const double term = (A * Fhi[n] - A * Flo[n]);
        du[n] = d == 0 ? term : du[n] + term;
      }
      if (flux) {
        double *Fo = flux + d * slab + c0;
#pragma unroll
        for (int n = 0; n < V; ++n) Fo[n * cs] = Flo[n];
        const bool last = (d == 0 && i == G.hi0) || (d == 1 && j == G.hi1) || (d == 2 && k == G.hi2);
        if (last)
#pragma unroll
          for (int n = 0; n < V; ++n) Fo[n * cs + st[d]] = Fhi[n];
      }
    }
    const double bdt = beta * dtvar[0];
    const double w2 = 1.0 - beta;
    double u1[V];
#pragma unroll
    for (int n = 0; n < V; ++n) {
      const double dudt = -du[n] / C.vol;
      const double avg = beta * U0[c0 + n * cs] + w2 * base[c0 + n * cs];
      u1[n] = 1.0 * avg + bdt * dudt;
      out[c0 + n * cs] = u1[n];
    }
    if (derived)
      derived[static_cast<std::int64_t>(b) * cs + (k * G.e1 + j) * G.e0 + i] =
          0.5 * u1[3] * (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    double s = fabs(u1[0]) / C.dx[0];
    if (G.ndim > 1) s = s + fabs(u1[1]) / C.dx[1];
    if (G.ndim > 2) s = s + fabs(u1[2]) / C.dx[2];
    m = 1.0 / s;
  }
  if (!partial) return;
  m = WaveMin(m);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = wmin[0];
#pragma unroll
    for (int w = 1; w < kWG / 64; ++w) v = DMin(v, wmin[w]);
    partial[L] = v;
  }
}

// ---- plane-tiled reconstruction + HLL (cubic blocks of N <= 16 cells per side) ----------
//
// ReconFluxKernel: one workgroup = PPW consecutive planes of one block for one direction. For
// each plane the WENO5Z (or linear) states of the plane's cells, extended by one cell along the
// direction, go to LDS; then the HLL flux of every face of the plane goes to the flux array
// [3][block][comp][brick] (x-face i at index i). The operations are those of ReconTaskKernel
// followed by HllFace, so the fluxes are the same bits, and the face states never reach HBM:
// the flux array is half their size and the update reads it without recomputing HLL on shared
// faces. PPW (planes per workgroup) changes only the work split and is picked by timing.
// PART selects the cells and faces of each plane row along the direction (ext index a of the
// N + 2 cells lo - 1 .. hi + 1; face f between ext cells f and f + 1):
//   0: every cell and face;
//   1: the cells whose WENO stencil reads no ghost cell (a in [3, N - 2]) and the faces
//      between two of them (f in [3, N - 3]): no dependence on the boundary exchange;
//   2: the rest (a in [0, 3] and [N - 2, N + 1], f in [0, 2] and [N - 2, N]); the cells 3 and
//      N - 2 are evaluated in both parts.
// The LDS holds the part's cells only: local cell q of a row is ext cell Cell(q), and face q
// lies between local cells Left(q) and Left(q) + 1.
template <int N, int PART>
struct PlaneSet {
  static constexpr int kCells = PART == 0 ? N + 2 : PART == 1 ? N - 4 : 8;
  static constexpr int kFaces = PART == 0 ? N + 1 : PART == 1 ? N - 5 : 6;
  __device__ static int Cell(int q) {
    return PART == 0 ? q : PART == 1 ? 3 + q : (q < 4 ? q : N - 6 + q);
  }
  __device__ static int Face(int q) {
    return PART == 0 ? q : PART == 1 ? 3 + q : (q < 3 ? q : N - 5 + q);
  }
  __device__ static int Left(int q) { return PART == 2 && q >= 3 ? q + 1 : q; }
};

template <int N, int NG, int V, int PART, int PPW>
__global__ void __launch_bounds__(kWG)
    ReconFluxKernel(const double *__restrict__ U, double *__restrict__ flux, int nblocks,
                    int recon) {
  constexpr int E = N + 2 * NG;  // brick extent
  using S = PlaneSet<N, PART>;
  constexpr int P = S::kCells * N;  // states of one plane per component
  constexpr int G = N / PPW;        // plane groups per (block, direction)
  __shared__ double sl[V * P], sr[V * P];
  // logical workgroup = (block * 3 + direction) * G + group; one block's planes are adjacent
  // and stay on one XCD
  const unsigned L = XcdRemap(blockIdx.x, gridDim.x);
  const int b = static_cast<int>(L / (3 * G)), r = static_cast<int>(L % (3 * G));
  const int d = r / G, plane0 = (r % G) * PPW;
  constexpr std::int64_t cs = static_cast<std::int64_t>(E) * E * E;
  const std::int64_t slab = cs * V * nblocks;
  const double *Ub = U + static_cast<std::int64_t>(b) * V * cs;
  const unsigned st = 8u * static_cast<unsigned>(d == 0 ? 1 : d == 1 ? E : E * E);
  for (int pp = 0; pp < PPW; ++pp) {
    const int plane = plane0 + pp;
    if (pp > 0) __syncthreads();  // the previous plane's faces have read the states
    // plane index of (local cell q along the direction, transverse tv): rows of kCells along
    // x for d = 0, else rows of N along x; consecutive threads take consecutive x
    for (int t = threadIdx.x; t < V * P; t += kWG) {
      const int n = t / P, u = t % P;
      int q, tv;
      if (d == 0) {
        q = u % S::kCells;
        tv = u / S::kCells;
      } else {
        tv = u % N;
        q = u / N;
      }
      const int a = S::Cell(q);
      int i, j, k;
      if (d == 0) {
        i = NG - 1 + a;
        j = NG + tv;
        k = NG + plane;
      } else if (d == 1) {
        i = NG + tv;
        j = NG - 1 + a;
        k = NG + plane;
      } else {
        i = NG + tv;
        j = NG + plane;
        k = NG - 1 + a;
      }
      const unsigned c = 8u * static_cast<unsigned>((k * E + j) * E + i);
      const double *qv = Ub + n * cs;
      double ql, qr;
      if (recon == 0)
        Weno5Z(LoadAt(qv, c - 2 * st), LoadAt(qv, c - st), LoadAt(qv, c), LoadAt(qv, c + st),
               LoadAt(qv, c + 2 * st), ql, qr);
      else
        LinearRecon(LoadAt(qv, c - st), LoadAt(qv, c), LoadAt(qv, c + st), ql, qr);
      sl[t] = ql;  // t = n * P + u and u is the plane index
      sr[t] = qr;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < S::kFaces * N; t += kWG) {
      int q, tv;
      if (d == 0) {
        q = t % S::kFaces;
        tv = t / S::kFaces;
      } else {
        tv = t % N;
        q = t / N;
      }
      const int f = S::Face(q), lq = S::Left(q);
      int pl, pr, i, j, k;
      if (d == 0) {  // face between ext cells f and f + 1 of row tv
        pl = tv * S::kCells + lq;
        pr = pl + 1;
        i = NG + f;
        j = NG + tv;
        k = NG + plane;
      } else {  // face between ext rows f and f + 1
        pl = lq * N + tv;
        pr = pl + N;
        i = NG + tv;
        j = d == 1 ? NG + f : NG + plane;
        k = d == 1 ? NG + plane : NG + f;
      }
      double Fv[V];
      HllFace<V>(sl + pl, sr + pr, P, d, Fv);
      double *fo = flux + d * slab + static_cast<std::int64_t>(b) * V * cs + (k * E + j) * E + i;
#pragma unroll
      for (int n = 0; n < V; ++n) fo[n * cs] = Fv[n];
    }
  }
}

// FluxUpdateKernel: FusedUpdateKernel reading stored face fluxes instead of recomputing HLL
// (flux divergence, RK averaging and update, derived field, CFL partial); same operations.
template <int V>
__global__ void __launch_bounds__(kWG)
    FluxUpdateKernel(const double *__restrict__ flux, const double *U0, const double *base,
                     double *out, double *__restrict__ derived, double *__restrict__ partial,
                     const BlockCoefs *__restrict__ coef, Geo G, int nblocks, double beta,
                     const double *__restrict__ dtvar) {
  __shared__ double wmin[kWG / 64];
  const int n0 = G.hi0 - G.lo0 + 1, n1 = G.hi1 - G.lo1 + 1, n2 = G.hi2 - G.lo2 + 1;
  const unsigned L = XcdRemap(blockIdx.x, gridDim.x);  // = block * tiles + tile
  const int tiles = (n0 * n1 * n2 + kWG - 1) / kWG;
  const int b = static_cast<int>(L / tiles);
  const int t = static_cast<int>(L % tiles) * kWG + threadIdx.x;
  double m = DBL_MAX;
  if (t < n0 * n1 * n2) {
    const int i = t % n0 + G.lo0;
    const int r = t / n0;
    const int j = r % n1 + G.lo1;
    const int k = r / n1 + G.lo2;
    const std::int64_t slab = G.cells * V * nblocks;
    const std::int64_t cs = G.cells;
    const std::int64_t c0 = static_cast<std::int64_t>(b) * V * cs + (k * G.e1 + j) * G.e0 + i;
    const BlockCoefs &C = coef[b];
    const std::int64_t st[3] = {1, G.e0, static_cast<std::int64_t>(G.e0) * G.e1};
    double du[V];
    for (int d = 0; d < G.ndim; ++d) {
      const double *Fd = flux + d * slab + c0;
      const double A = C.area[d];
#pragma unroll
      for (int n = 0; n < V; ++n) {
        const double term = (A * Fd[n * cs + st[d]] - A * Fd[n * cs]);
        du[n] = d == 0 ? term : du[n] + term;
      }
    }
    const double bdt = beta * dtvar[0];
    const double w2 = 1.0 - beta;
    double u1[V];
#pragma unroll
    for (int n = 0; n < V; ++n) {
      const double dudt = -du[n] / C.vol;
      const double avg = beta * U0[c0 + n * cs] + w2 * base[c0 + n * cs];
      u1[n] = 1.0 * avg + bdt * dudt;
      out[c0 + n * cs] = u1[n];
    }
    if (derived)
      derived[static_cast<std::int64_t>(b) * cs + (k * G.e1 + j) * G.e0 + i] =
          0.5 * u1[3] * (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    double sp = fabs(u1[0]) / C.dx[0];
    if (G.ndim > 1) sp = sp + fabs(u1[1]) / C.dx[1];
    if (G.ndim > 2) sp = sp + fabs(u1[2]) / C.dx[2];
    m = 1.0 / sp;
  }
  if (!partial) return;
  m = WaveMin(m);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = wmin[0];
#pragma unroll
    for (int w = 1; w < kWG / 64; ++w) v = DMin(v, wmin[w]);
    partial[L] = v;
  }
}

// RestrictAverage::Do for face element F(D+1) (pr_ops.hpp:105-164): area-weighted average of
// the fine face fluxes fqp (p along the lower transverse dimension t1, q along t2) sharing one
// coarse face, in the reference's fixed summation order; faces beyond the active dimensions
// carry zero weight and a zero term, as in the reference loop.
template <int D>
__device__ __forceinline__ double RestrictFaceValue(double area, double f00, double f01,
                                                    double f10, double f11, int s1, int s2) {
  double T[2][2][2], W[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) T[a][b][c] = W[a][b][c] = 0.0;
  const double f[2][2] = {{f00, f01}, {f10, f11}};
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (p >= s1 || q >= s2) continue;
      // [ok][oj][oi]: D=0 -> (t2 = z, t1 = y); D=1 -> (t2 = z, t1 = x); D=2 -> (t2 = y, t1 = x)
      double &t = D == 0 ? T[q][p][0] : D == 1 ? T[q][0][p] : T[0][q][p];
      double &w = D == 0 ? W[q][p][0] : D == 1 ? W[q][0][p] : W[0][q][p];
      w = area;
      t = w * f[q][p];
    }
  const double tvol = ((W[0][0][0] + W[0][1][0]) + (W[0][0][1] + W[0][1][1])) +
                      ((W[1][0][0] + W[1][1][0]) + (W[1][0][1] + W[1][1][1]));
  return (((T[0][0][0] + T[0][1][0]) + (T[0][0][1] + T[0][1][1])) +
          ((T[1][0][0] + T[1][1][0]) + (T[1][0][1] + T[1][1][1]))) /
         tvol;
}

__device__ __forceinline__ double RestrictFace(int d, double area, double f00, double f01,
                                               double f10, double f11, int s1, int s2) {
  if (d == 0) return RestrictFaceValue<0>(area, f00, f01, f10, f11, s1, s2);
  if (d == 1) return RestrictFaceValue<1>(area, f00, f01, f10, f11, s1, s2);
  return RestrictFaceValue<2>(area, f00, f01, f10, f11, s1, s2);
}

// Brick offset of fine face (p, q) of patch cell (a, c): the face index along d, and transverse
// fine cells step * (a, c) + (p, q) from the patch start.
__device__ __forceinline__ std::int64_t PatchFineCell(const FluxPatch &P, const Geo &G, int a,
                                                      int c, int p, int q) {
  const int u = P.ft[0] + P.step[0] * a + p, v = P.ft[1] + P.step[1] * c + q;
  const int i = P.d == 0 ? P.fface : u;
  const int j = P.d == 1 ? P.fface : (P.d == 0 ? u : v);
  const int k = P.d == 2 ? P.fface : v;
  return (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
}

__device__ __forceinline__ std::int64_t PatchCoarseCell(const FluxPatch &P, const Geo &G, int a,
                                                        int c) {
  const int u = P.ct[0] + a, v = P.ct[1] + c;
  const int i = P.d == 0 ? P.cface : u;
  const int j = P.d == 1 ? P.cface : (P.d == 0 ? u : v);
  const int k = P.d == 2 ? P.cface : v;
  return (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i;
}

// flxcor_send of the fused stage: one thread per coarse face cell of a patch evaluates the HLL
// fluxes of its 2x2 fine faces from the fine block's face states (the values CalculateFluxes
// would store there) and writes their restriction to the correction store.
template <int V>
__global__ void __launch_bounds__(64)
    FlxcorFaceKernel(const FluxPatch *__restrict__ patches, const double *__restrict__ face,
                     double *__restrict__ store, Geo G, int nblocks) {
  const FluxPatch &P = patches[blockIdx.x];
  const int t = blockIdx.y * 64 + threadIdx.x;
  if (t >= P.n[0] * P.n[1]) return;
  const int a = t % P.n[0], c = t / P.n[0];
  const std::int64_t cs = G.cells, slab = G.cells * V * nblocks;
  const std::int64_t base = static_cast<std::int64_t>(P.fine) * V * cs;
  double F[2][2][V];
#pragma unroll
  for (int q = 0; q < 2; ++q)
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (p >= P.step[0] || q >= P.step[1]) {
#pragma unroll
        for (int n = 0; n < V; ++n) F[q][p][n] = 0.0;
        continue;
      }
      const std::int64_t o = base + PatchFineCell(P, G, a, c, p, q);
      HllFace<V>(face + (2 * P.d) * slab + o, face + (2 * P.d + 1) * slab + o, cs, P.d,
                 F[q][p]);
    }
  const std::int64_t dst = P.dst + a + static_cast<std::int64_t>(c) * P.rows;
#pragma unroll
  for (int n = 0; n < V; ++n)
    store[dst + static_cast<std::int64_t>(n) * P.comp] =
        RestrictFace(P.d, P.area, F[0][0][n], F[0][1][n], F[1][0][n], F[1][1][n], P.step[0],
                     P.step[1]);
}

// SendBoundBufs<flxcor_send> over stored face fluxes [dir][block][comp][brick].
__global__ void __launch_bounds__(64)
    FlxcorSendKernel(const FluxPatch *__restrict__ patches, const double *__restrict__ flux,
                     double *__restrict__ store, Geo G, int nblocks, int nc) {
  const FluxPatch &P = patches[blockIdx.x];
  const int t = blockIdx.y * 64 + threadIdx.x;
  if (t >= P.n[0] * P.n[1]) return;
  const int a = t % P.n[0], c = t / P.n[0];
  const std::int64_t cs = G.cells, slab = G.cells * nc * nblocks;
  const double *F = flux + P.d * slab + static_cast<std::int64_t>(P.fine) * nc * cs;
  const std::int64_t o00 = PatchFineCell(P, G, a, c, 0, 0);
  const std::int64_t o01 = P.step[0] > 1 ? PatchFineCell(P, G, a, c, 1, 0) : o00;
  const std::int64_t o10 = P.step[1] > 1 ? PatchFineCell(P, G, a, c, 0, 1) : o00;
  const std::int64_t o11 =
      P.step[0] > 1 && P.step[1] > 1 ? PatchFineCell(P, G, a, c, 1, 1) : o00;
  const std::int64_t dst = P.dst + a + static_cast<std::int64_t>(c) * P.rows;
  for (int n = 0; n < nc; ++n) {
    const double *Fn = F + n * cs;
    store[dst + static_cast<std::int64_t>(n) * P.comp] =
        RestrictFace(P.d, P.area, Fn[o00], Fn[o01], Fn[o10], Fn[o11], P.step[0], P.step[1]);
  }
}

// SetBounds<flxcor_recv>: the coarse block's stored fluxes on the patch's part of its face.
__global__ void __launch_bounds__(64)
    FlxcorSetKernel(const FluxPatch *__restrict__ patches, const double *__restrict__ store,
                    double *__restrict__ flux, Geo G, int nblocks, int nc) {
  const FluxPatch &P = patches[blockIdx.x];
  const int t = blockIdx.y * 64 + threadIdx.x;
  if (t >= P.n[0] * P.n[1]) return;
  const int a = t % P.n[0], c = t / P.n[0];
  const std::int64_t cs = G.cells, slab = G.cells * nc * nblocks;
  double *F = flux + P.d * slab + static_cast<std::int64_t>(P.coarse) * nc * cs +
              PatchCoarseCell(P, G, a, c);
  const std::int64_t src = P.dst + a + static_cast<std::int64_t>(c) * P.rows;
  for (int n = 0; n < nc; ++n) F[n * cs] = store[src + static_cast<std::int64_t>(n) * P.comp];
}

// Left and right states of one cell from its stencil along a direction (byte offsets c of the
// cell and st of one step): WENO5Z or the MC-limited linear reconstruction.
__device__ __forceinline__ void CellStates(const double *q, unsigned c, unsigned st, int recon,
                                           double &ql, double &qr) {
  if (recon == 0)
    Weno5Z(LoadAt(q, c - 2 * st), LoadAt(q, c - st), LoadAt(q, c), LoadAt(q, c + st),
           LoadAt(q, c + 2 * st), ql, qr);
  else
    LinearRecon(LoadAt(q, c - st), LoadAt(q, c), LoadAt(q, c + st), ql, qr);
}

// flxcor_send of the block-sweep stage: four adjacent threads per coarse face cell of a patch,
// one per fine face (p, q), evaluate the HLL flux of their face straight from the fine block's
// state (the left state of face f is ql of cell f - 1, the right state qr of cell f, as
// CalculateFluxes stores them); the first of the four gathers the other three by lane
// shuffles and writes their restriction to the correction store.
template <int V>
__global__ void __launch_bounds__(64)
    FlxcorDirectKernel(const FluxPatch *__restrict__ patches, const double *__restrict__ U,
                       double *__restrict__ store, Geo G, int recon) {
  const FluxPatch &P = patches[blockIdx.x];
  const int t = blockIdx.y * 64 + threadIdx.x;
  const int cell = t >> 2, sub = t & 3, p = sub & 1, q = sub >> 1;
  const bool live = cell < P.n[0] * P.n[1];
  const int a = live ? cell % P.n[0] : 0, c = live ? cell / P.n[0] : 0;
  const std::int64_t cs = G.cells;
  const double *Ub = U + static_cast<std::int64_t>(P.fine) * V * cs;
  const unsigned st = 8u * static_cast<unsigned>(P.d == 0 ? 1 : P.d == 1 ? G.e0 : G.e0 * G.e1);
  double F[V];
#pragma unroll
  for (int n = 0; n < V; ++n) F[n] = 0.0;
  if (live && p < P.step[0] && q < P.step[1]) {
    const unsigned o = 8u * static_cast<unsigned>(PatchFineCell(P, G, a, c, p, q));
    double l[V], r[V];
#pragma unroll
    for (int n = 0; n < V; ++n) {
      double x, y;
      CellStates(Ub + n * cs, o - st, st, recon, l[n], x);
      CellStates(Ub + n * cs, o, st, recon, y, r[n]);
    }
    HllGen<V>([&](int n) { return l[n]; }, [&](int n) { return r[n]; }, P.d, F);
  }
  const int lane0 = static_cast<int>(threadIdx.x) & ~3;
  const std::int64_t dst = P.dst + a + static_cast<std::int64_t>(c) * P.rows;
#pragma unroll
  for (int n = 0; n < V; ++n) {
    const double f00 = __shfl(F[n], lane0, 64), f01 = __shfl(F[n], lane0 + 1, 64);
    const double f10 = __shfl(F[n], lane0 + 2, 64), f11 = __shfl(F[n], lane0 + 3, 64);
    if (live && sub == 0)
      store[dst + static_cast<std::int64_t>(n) * P.comp] =
          RestrictFace(P.d, P.area, f00, f01, f10, f11, P.step[0], P.step[1]);
  }
}

// ---- block sweep: reconstruction, HLL, flux divergence and RK update in one launch ----------
//
// One thread per (i, j) column of a cubic block of N cells per side, marching k over interior
// planes. Per plane: the x states of the plane's rows (extended by one cell each way) go to
// LDS, the HLL flux of every x face of the plane is evaluated once into LDS (each thread the
// high face of its cell, the first N threads the low face of one row each), and each thread
// forms its cell's x difference; then the same for y; then the z states of the plane above,
// whose qr closes the high z face (the low z face of plane k is the high z face of plane
// k - 1, carried in registers). The flux divergence sums x, y, z in the order of
// FluxDivHelper (update.hpp:43-57), then the RK averaging and update, the derived field and
// the CFL minimum follow as in FluxUpdateKernel. Every face flux is the same expression as in
// the other forms, so the stage output is the same bits, and no flux or face state reaches
// HBM. On
// multilevel meshes, faces with a finer neighbour take the restricted fine fluxes from the
// correction store (fover / fcor as in FusedUpdateKernel).
// The grid is persistent: workgroup g takes planes [g T / G, (g + 1) T / G) of the
// T = nblocks * N planes in block order, so every workgroup does the same number of planes
// (a segment that starts inside a block first rebuilds the z state below its first plane).
// partial[g] is the workgroup's CFL minimum.

template <int N, int NG, int V, int MINW>
__global__ void __launch_bounds__(N * N, MINW)
    BlockSweepKernel(const double *__restrict__ U, const double *base, double *out,
                     double *__restrict__ derived, double *__restrict__ partial,
                     const BlockCoefs *__restrict__ coef, int nblocks, double beta,
                     const double *__restrict__ dtvar, int recon, const int *__restrict__ fover,
                     const double *__restrict__ fcor, const int *__restrict__ push) {
  static_assert(N >= 2 * NG, "a cell lies in at most one boundary strip per dimension");
  constexpr int E = N + 2 * NG;
  constexpr int NT = N * N;
  constexpr int R = N + 2;  // states per row: ext cells lo - 1 .. hi + 1
  constexpr int P = R * N;  // states of one plane per component
  constexpr int FP = (N + 1) * N;  // faces of one plane per component and direction
  __shared__ double sl[V * P], sr[V * P];
  __shared__ double fx[V * FP];  // x (then y) face fluxes of the plane, each face once
  __shared__ double wmin[(NT + 63) / 64];
  const int tx = threadIdx.x % N, ty = threadIdx.x / N;
  constexpr std::int64_t cs = static_cast<std::int64_t>(E) * E * E;
  const double bdt = beta * dtvar[0];
  const double w2 = 1.0 - beta;
  constexpr int comp = N * N;  // correction-store stride between components
  constexpr unsigned sx = 8u, sy = 8u * E, sz = 8u * E * E;
  double m = DBL_MAX;
  const std::int64_t T = static_cast<std::int64_t>(nblocks) * N;
  std::int64_t pnext = T * blockIdx.x / gridDim.x;
  const std::int64_t pend = T * (blockIdx.x + 1) / gridDim.x;
  while (pnext < pend) {
    const int b = static_cast<int>(pnext / N);
    const int k0 = static_cast<int>(pnext % N);
    const int k1 = static_cast<int>(pend - pnext < N - k0 ? k0 + (pend - pnext) : N);
    pnext += k1 - k0;
    const double *Ub = U + static_cast<std::int64_t>(b) * V * cs;
    const BlockCoefs &C = coef[b];
    const int ox0 = fover ? fover[6 * b + 0] : -1, ox1 = fover ? fover[6 * b + 1] : -1;
    const int oy0 = fover ? fover[6 * b + 2] : -1, oy1 = fover ? fover[6 * b + 3] : -1;
    const int oz0 = fover ? fover[6 * b + 4] : -1, oz1 = fover ? fover[6 * b + 5] : -1;
    // z states of one plane of cells kc for every column, [n][j][i] (a flat loop like x and y)
    auto zstates = [&](int kc) {
      for (int t = threadIdx.x; t < V * NT; t += NT) {
        const int n = t / NT, u = t % NT;
        const unsigned c = 8u * static_cast<unsigned>((kc * E + NG + u / N) * E + NG + u % N);
        CellStates(Ub + n * cs, c, sz, recon, sl[n * P + u], sr[n * P + u]);
      }
    };
    // z faces: zl = ql of the cell below the current plane, zf = flux on the current low z face
    double zl[V], zf[V];
    __syncthreads();  // the previous segment's states have been read
    zstates(NG + k0 - 1);
    __syncthreads();
#pragma unroll
    for (int n = 0; n < V; ++n) zl[n] = sl[n * P + threadIdx.x];
    __syncthreads();
    zstates(NG + k0);
    __syncthreads();
    HllGen<V>([&](int n) { return zl[n]; }, [&](int n) { return sr[n * P + threadIdx.x]; }, 2, zf);
    if (k0 == 0 && oz0 >= 0)
#pragma unroll
      for (int n = 0; n < V; ++n) zf[n] = fcor[oz0 + tx + ty * N + n * comp];
#pragma unroll
    for (int n = 0; n < V; ++n) zl[n] = sl[n * P + threadIdx.x];
    for (int k = k0; k < k1; ++k) {
      const int kk = NG + k;
      double du[V];
      // ---- x: states of ext cells lo - 1 .. hi + 1 of every row of the plane, [n][row][a]
      __syncthreads();  // the previous states have been read
      for (int t = threadIdx.x; t < V * P; t += NT) {
        const int n = t / P, u = t % P;
        const int a = u % R, row = u / R;
        const unsigned c = 8u * static_cast<unsigned>((kk * E + NG + row) * E + NG - 1 + a);
        CellStates(Ub + n * cs, c, sx, recon, sl[t], sr[t]);
      }
      __syncthreads();
      {
        // every x face once: each thread the high face of its cell, threads 0 .. N - 1 also
        // the low face of row threadIdx.x; fluxes to LDS as [n][row][face]
        double F[V];
        const double *L = sl + ty * R + tx, *Rr = sr + ty * R + tx;
        HllGen<V>([&](int n) { return L[n * P + 1]; }, [&](int n) { return Rr[n * P + 2]; }, 0,
                  F);
        if (tx == N - 1 && ox1 >= 0)
#pragma unroll
          for (int n = 0; n < V; ++n) F[n] = fcor[ox1 + ty + k * N + n * comp];
#pragma unroll
        for (int n = 0; n < V; ++n) fx[n * FP + ty * (N + 1) + tx + 1] = F[n];
        if (static_cast<int>(threadIdx.x) < N) {
          const int row = threadIdx.x;
          HllGen<V>([&](int n) { return sl[n * P + row * R]; },
                    [&](int n) { return sr[n * P + row * R + 1]; }, 0, F);
          if (ox0 >= 0)
#pragma unroll
            for (int n = 0; n < V; ++n) F[n] = fcor[ox0 + row + k * N + n * comp];
#pragma unroll
          for (int n = 0; n < V; ++n) fx[n * FP + row * (N + 1)] = F[n];
        }
      }
      __syncthreads();  // the x fluxes are in LDS and the x states have been read
      {
        // du = A F(hi) - A F(lo)
        const double A = C.area[0];
        const double *f = fx + ty * (N + 1) + tx;
#pragma unroll
        for (int n = 0; n < V; ++n) du[n] = (A * f[n * FP + 1] - A * f[n * FP]);
      }
      // ---- y: states of ext cells lo - 1 .. hi + 1 of every column, [n][a][i]
      for (int t = threadIdx.x; t < V * P; t += NT) {
        const int n = t / P, u = t % P;
        const int i = u % N, a = u / N;
        const unsigned c = 8u * static_cast<unsigned>((kk * E + NG - 1 + a) * E + NG + i);
        CellStates(Ub + n * cs, c, sy, recon, sl[t], sr[t]);
      }
      __syncthreads();  // also: every thread has read its x fluxes
      {
        // every y face once, as [n][face][i]
        double F[V];
        const double *L = sl + ty * N + tx, *Rr = sr + ty * N + tx;
        HllGen<V>([&](int n) { return L[n * P + N]; }, [&](int n) { return Rr[n * P + 2 * N]; },
                  1, F);
        if (ty == N - 1 && oy1 >= 0)
#pragma unroll
          for (int n = 0; n < V; ++n) F[n] = fcor[oy1 + tx + k * N + n * comp];
#pragma unroll
        for (int n = 0; n < V; ++n) fx[n * FP + (ty + 1) * N + tx] = F[n];
        if (static_cast<int>(threadIdx.x) < N) {
          const int col = threadIdx.x;
          HllGen<V>([&](int n) { return sl[n * P + col]; },
                    [&](int n) { return sr[n * P + N + col]; }, 1, F);
          if (oy0 >= 0)
#pragma unroll
            for (int n = 0; n < V; ++n) F[n] = fcor[oy0 + col + k * N + n * comp];
#pragma unroll
          for (int n = 0; n < V; ++n) fx[n * FP + col] = F[n];
        }
      }
      __syncthreads();  // the y fluxes are in LDS and the y states have been read
      {
        const double A = C.area[1];
        const double *f = fx + ty * N + tx;
#pragma unroll
        for (int n = 0; n < V; ++n) du[n] = du[n] + (A * f[n * FP + N] - A * f[n * FP]);
      }
      // ---- z: qr of the cell above this plane gives the high z face; its ql moves down
      zstates(kk + 1);
      __syncthreads();  // also: every thread has read its y fluxes
      {
        double Fhi[V];
        HllGen<V>([&](int n) { return zl[n]; }, [&](int n) { return sr[n * P + threadIdx.x]; }, 2,
                  Fhi);
        if (k == N - 1 && oz1 >= 0)
#pragma unroll
          for (int n = 0; n < V; ++n) Fhi[n] = fcor[oz1 + tx + ty * N + n * comp];
        const double A = C.area[2];
#pragma unroll
        for (int n = 0; n < V; ++n) {
          du[n] = du[n] + (A * Fhi[n] - A * zf[n]);
          zf[n] = Fhi[n];
          zl[n] = sl[n * P + threadIdx.x];
        }
      }
      // ---- divergence, RK averaging and update, derived field, CFL
      const std::int64_t cell = (static_cast<std::int64_t>(kk) * E + NG + ty) * E + NG + tx;
      const std::int64_t c0 = static_cast<std::int64_t>(b) * V * cs + cell;
      double u1[V];
#pragma unroll
      for (int n = 0; n < V; ++n) {
        const double dudt = -du[n] / C.vol;
        const double avg = beta * U[c0 + n * cs] + w2 * base[c0 + n * cs];
        u1[n] = 1.0 * avg + bdt * dudt;
        out[c0 + n * cs] = u1[n];
      }
      if (push) {
        // the exchange's same-level copies: a cell within NG of a face is a ghost cell of the
        // same-level neighbours across that face, edge or corner (push[27 b + direction]),
        // at its index shifted by N along each crossed dimension
        const int lx = tx < NG ? -1 : (tx >= N - NG ? 1 : 0);
        const int ly = ty < NG ? -1 : (ty >= N - NG ? 1 : 0);
        const int lz = k < NG ? -1 : (k >= N - NG ? 1 : 0);
        for (int az = 0; az < (lz ? 2 : 1); ++az)
          for (int ay = 0; ay < (ly ? 2 : 1); ++ay)
            for (int ax = 0; ax < (lx ? 2 : 1); ++ax) {
              const int dx = ax ? lx : 0, dy = ay ? ly : 0, dz = az ? lz : 0;
              if (!(dx | dy | dz)) continue;
              const int g = push[27 * b + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)];
              if (g < 0) continue;
              const std::int64_t d0 =
                  static_cast<std::int64_t>(g) * V * cs +
                  (static_cast<std::int64_t>(kk - dz * N) * E + NG + ty - dy * N) * E + NG + tx -
                  dx * N;
#pragma unroll
              for (int n = 0; n < V; ++n) out[d0 + n * cs] = u1[n];
            }
      }
      if (derived)
        derived[static_cast<std::int64_t>(b) * cs + cell] =
            0.5 * u1[3] * (u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
      double sp = fabs(u1[0]) / C.dx[0];
      sp = sp + fabs(u1[1]) / C.dx[1];
      sp = sp + fabs(u1[2]) / C.dx[2];
      m = DMin(m, 1.0 / sp);
    }
  }
  if (!partial) return;
  m = WaveMin(m);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = wmin[0];
#pragma unroll
    for (int w = 1; w < (NT + 63) / 64; ++w) v = DMin(v, wmin[w]);
    partial[blockIdx.x] = v;
  }
}

// burgers_package::MassHistory (benchmarks/burgers/burgers_package.cpp:406-439) for the 8
// octants of the mesh at once: sum over interior cells and components of mask * q * q * w,
// w = cell volume / (mesh volume + 1e-20), mask from the inclusive cell-centre test.
// Per-workgroup partials [workgroup][8]; the order of the sum is this build's own (the
// reference's Kokkos::Sum order is not fixed either).
struct Octants {
  double lo[8][3], hi[8][3];
};
__global__ void __launch_bounds__(kWG)
    MassHistoryKernel(const double *__restrict__ U, const DeviceBlockGeom *__restrict__ geom,
                      const BlockCoefs *__restrict__ coef, Geo G, int nblocks, Octants O,
                      double mesh_vol, double *__restrict__ partial) {
  __shared__ double acc[kWG / 64][8];
  const int n0 = G.hi0 - G.lo0 + 1, n1 = G.hi1 - G.lo1 + 1, n2 = G.hi2 - G.lo2 + 1;
  const std::int64_t per = static_cast<std::int64_t>(n0) * n1 * n2;
  const std::int64_t total = per * G.V * nblocks;
  double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<std::int64_t>(gridDim.x) * kWG) {
    std::int64_t r = t;
    const int i = static_cast<int>(r % n0) + G.lo0;
    r /= n0;
    const int j = static_cast<int>(r % n1) + G.lo1;
    r /= n1;
    const int k = static_cast<int>(r % n2) + G.lo2;
    r /= n2;  // block * V + comp
    const int b = static_cast<int>(r / G.V);
    const DeviceBlockGeom &g = geom[b];
    const double w = coef[b].vol / (mesh_vol + 1e-20);
    const double x1 = g.x0[0] + (i + 0.5) * g.dx[0];
    const double x2 = g.x0[1] + (j + 0.5) * g.dx[1];
    const double x3 = g.x0[2] + (k + 0.5) * g.dx[2];
    const double q = U[r * G.cells + (static_cast<std::int64_t>(k) * G.e1 + j) * G.e0 + i];
#pragma unroll
    for (int o = 0; o < 8; ++o) {
      const double mask = (O.lo[o][0] <= x1) && (x1 <= O.hi[o][0]) && (O.lo[o][1] <= x2) &&
                          (x2 <= O.hi[o][1]) && (O.lo[o][2] <= x3) && (x3 <= O.hi[o][2]);
      s[o] += mask * q * q * w;
    }
  }
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    double v = s[o];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) acc[threadIdx.x >> 6][o] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double v = 0.0;
    for (int w2 = 0; w2 < kWG / 64; ++w2) v += acc[w2][threadIdx.x];
    partial[8 * blockIdx.x + threadIdx.x] = v;
  }
}

__global__ void __launch_bounds__(64)
    SumPartials8Kernel(const double *__restrict__ partial, int n, double *__restrict__ out) {
  if (threadIdx.x >= 8) return;
  double v = 0.0;
  for (int q = 0; q < n; ++q) v += partial[8 * q + threadIdx.x];
  out[threadIdx.x] = v;
}

// Final CFL minimum and EvolutionDriver::SetGlobalTimeStep on device (driver.cpp:210-235):
// dtvar = {dt, time, estimate, dt of the last cycle}. advance 1 (end of a cycle): time += dt,
// the cycle's dt is kept, dt = min(2 dt, cfl * min); advance 0 (start): dt = min(2 dt, est)
// with dt still at its initial maximum; advance 2 (after a remesh): the estimate of the new
// blocks replaces the old one, dt = min(2 dt_last, cfl * min).
// One workgroup of kClockWG threads: strided partial minima, wavefront shuffles, then LDS.
constexpr int kClockWG = 1024;
__global__ void __launch_bounds__(kClockWG)
    ClockKernel(const double *__restrict__ partial, int n, double cfl, double *dtvar,
                int advance) {
  __shared__ double wmin[kClockWG / 64];
  double m = DBL_MAX;
  for (int t = threadIdx.x; t < n; t += kClockWG) m = DMin(m, partial[t]);
  m = WaveMin(m);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = wmin[0];
    for (int w = 1; w < kClockWG / 64; ++w) m = DMin(m, wmin[w]);
    const double est = cfl * m;
    double dt = advance == 2 ? dtvar[3] : dtvar[0];
    if (advance == 1) {
      dtvar[1] = dtvar[1] + dt;
      dtvar[3] = dt;
    }
    if (dt < 0.1 * DBL_MAX) dt *= 2.0;
    dtvar[0] = DMin(dt, est);
    dtvar[2] = est;
  }
}

int Grid(std::int64_t total) {
  const std::int64_t want = (total + kWG - 1) / kWG;
  return static_cast<int>(want < 8192 ? want : 8192);
}

Geo MakeGeo(const Topology &t, int V) {
  Geo g;
  g.e0 = t.fine.ent(0);
  g.e1 = t.fine.ent(1);
  g.e2 = t.fine.ent(2);
  g.lo0 = t.fine.lo(0);
  g.lo1 = t.fine.lo(1);
  g.lo2 = t.fine.lo(2);
  g.hi0 = t.fine.hi(0);
  g.hi1 = t.fine.hi(1);
  g.hi2 = t.fine.hi(2);
  g.V = V;
  g.ndim = t.ndim;
  g.cells = t.fine.cells();
  return g;
}

}  // namespace

pth_status BurgersRecon(const Topology &t, int V, const double *U, double *face, int recon,
                        hipStream_t s) {
  ProfScope prof("recon", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const int dk = G.ndim > 2 ? 1 : 0, dj = G.ndim > 1 ? 1 : 0;
  const std::int64_t total = static_cast<std::int64_t>(G.hi0 - G.lo0 + 3) *
                             (G.hi1 - G.lo1 + 1 + 2 * dj) * (G.hi2 - G.lo2 + 1 + 2 * dk) * V * nb;
  hipLaunchKernelGGL(ReconKernel, dim3(Grid(total)), dim3(kWG), 0, s, U, face, G, nb, recon);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersFlux(const Topology &t, int V, const double *face, double *flux,
                       hipStream_t s) {
  ProfScope prof("hll_flux", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const int dk = G.ndim > 2 ? 1 : 0, dj = G.ndim > 1 ? 1 : 0;
  const std::int64_t total = static_cast<std::int64_t>(G.hi0 - G.lo0 + 2) *
                             (G.hi1 - G.lo1 + 1 + dj) * (G.hi2 - G.lo2 + 1 + dk) * nb;
  hipLaunchKernelGGL(FluxKernel, dim3(Grid(total)), dim3(kWG), 0, s, face, flux, G, nb);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status FluxDivergence(const Topology &t, int V, const double *flux, double *dudt,
                          const BlockCoefs *coef, hipStream_t s) {
  ProfScope prof("flux_div", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const std::int64_t total =
      static_cast<std::int64_t>(t.fine.n[0]) * t.fine.n[1] * t.fine.n[2] * V * nb;
  hipLaunchKernelGGL(FluxDivKernel, dim3(Grid(total)), dim3(kWG), 0, s, flux, dudt, coef, G, nb);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status WeightedSum(const double *x, const double *y, double w1, double w2, double *z,
                       std::int64_t n, hipStream_t s) {
  if (n == 0) return PTH_COMPLETE;
  ProfScope prof("weighted_sum", s);
  hipLaunchKernelGGL(WsumKernel, dim3(Grid(n)), dim3(kWG), 0, s, x, y, w1, w2, z, n);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersDerived(const Topology &t, int V, const double *U, double *out,
                          hipStream_t s) {
  ProfScope prof("derived", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const std::int64_t total =
      static_cast<std::int64_t>(t.fine.n[0]) * t.fine.n[1] * t.fine.n[2] * nb;
  hipLaunchKernelGGL(DerivedKernel, dim3(Grid(total)), dim3(kWG), 0, s, U, out, G, nb);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersReconTasks(const Topology &t, int V, const double *U, double *face, int recon,
                             hipStream_t s) {
  ProfScope prof("recon_tasks", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const int n0 = t.fine.n[0], n1 = t.fine.n[1], n2 = t.fine.n[2];
  int ntask = (n0 + 2) * n1 * n2;
  if (G.ndim > 1) ntask = std::max(ntask, n0 * (n1 + 2) * n2);
  if (G.ndim > 2) ntask = std::max(ntask, n0 * n1 * (n2 + 2));
  const unsigned per_brick = static_cast<unsigned>(((ntask + kWG - 1) / kWG) * G.ndim);
  const dim3 grid(per_brick * nb * V);
  const bool cubic = G.ndim == 3 && n0 == n1 && n1 == n2 && t.fine.g[0] == t.fine.g[1] &&
                     t.fine.g[1] == t.fine.g[2];
  const int g = t.fine.g[0];
  auto run = [&](auto kernel) {
    hipLaunchKernelGGL(kernel, grid, dim3(kWG), 0, s, U, face, G, nb, recon, per_brick);
  };
  if (cubic && n0 == 16 && g == 4)
    run(ReconTaskKernel<16, 4>);
  else if (cubic && n0 == 16 && g == 3)
    run(ReconTaskKernel<16, 3>);
  else if (cubic && n0 == 32 && g == 4)
    run(ReconTaskKernel<32, 4>);
  else if (cubic && n0 == 8 && g == 4)
    run(ReconTaskKernel<8, 4>);
  else
    run(ReconTaskKernel<0, 0>);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

int FusedUpdatePartials(const Topology &t) {
  const int cells = t.fine.n[0] * t.fine.n[1] * t.fine.n[2];
  return ((cells + kWG - 1) / kWG) * static_cast<int>(t.blocks.size());
}

pth_status BurgersFusedUpdate(const Topology &t, int V, const double *face, const double *U0,
                              const double *base, double *out, double *derived, double *flux,
                              double *partial, const BlockCoefs *coef, double beta,
                              const double *dtvar, const FluxCorrection *fc, hipStream_t s) {
  if (V > kMaxV) {
    SetError("fused Burgers update supports at most 16 components");
    return PTH_FAIL;
  }
  if (V < 4) {
    SetError("fused Burgers update needs at least 4 components");
    return PTH_FAIL;
  }
  ProfScope prof("fused_update", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const int cells = t.fine.n[0] * t.fine.n[1] * t.fine.n[2];
  const dim3 grid(((cells + kWG - 1) / kWG) * nb);
  const bool over = fc && !fc->host.empty();
  const int *fover = over ? fc->dev_face : nullptr;
  const double *fcor = over ? fc->store : nullptr;
#define PTH_FUSED_CASE(NV)                                                                     \
  case NV:                                                                                     \
    hipLaunchKernelGGL(FusedUpdateKernel<NV>, grid, dim3(kWG), 0, s, face, U0, base, out,      \
                       derived, flux, partial, coef, G, nb, beta, dtvar, fover, fcor);         \
    break;
  switch (V) {
    PTH_FUSED_CASE(4)
    PTH_FUSED_CASE(5)
    PTH_FUSED_CASE(6)
    PTH_FUSED_CASE(7)
    PTH_FUSED_CASE(8)
    PTH_FUSED_CASE(9)
    PTH_FUSED_CASE(10)
    PTH_FUSED_CASE(11)
    PTH_FUSED_CASE(12)
    PTH_FUSED_CASE(13)
    PTH_FUSED_CASE(14)
    PTH_FUSED_CASE(15)
    PTH_FUSED_CASE(16)
  }
#undef PTH_FUSED_CASE
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status FlxcorFromFaces(const Topology &t, int V, const FluxCorrection &fc,
                          const double *face, hipStream_t s) {
  if (fc.host.empty()) return PTH_COMPLETE;
  if (V > kMaxV || V < 4 || fc.ncomp != V) {
    SetError("flux correction from face states: component count mismatch");
    return PTH_FAIL;
  }
  ProfScope prof("flxcor_faces", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const dim3 grid(static_cast<unsigned>(fc.host.size()), (fc.max_cells + 63) / 64);
#define PTH_FLX_CASE(NV)                                                                     \
  case NV:                                                                                   \
    hipLaunchKernelGGL(FlxcorFaceKernel<NV>, grid, dim3(64), 0, s, fc.dev, face, fc.store, G, \
                       nb);                                                                  \
    break;
  switch (V) {
    PTH_FLX_CASE(4)
    PTH_FLX_CASE(5)
    PTH_FLX_CASE(6)
    PTH_FLX_CASE(7)
    PTH_FLX_CASE(8)
    PTH_FLX_CASE(9)
    PTH_FLX_CASE(10)
    PTH_FLX_CASE(11)
    PTH_FLX_CASE(12)
    PTH_FLX_CASE(13)
    PTH_FLX_CASE(14)
    PTH_FLX_CASE(15)
    PTH_FLX_CASE(16)
  }
#undef PTH_FLX_CASE
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status FlxcorSend(const Topology &t, const FluxCorrection &fc, const double *flux,
                      hipStream_t s) {
  if (fc.host.empty()) return PTH_COMPLETE;
  ProfScope prof("flxcor_send", s);
  const Geo G = MakeGeo(t, fc.ncomp);
  const int nb = static_cast<int>(t.blocks.size());
  const dim3 grid(static_cast<unsigned>(fc.host.size()), (fc.max_cells + 63) / 64);
  hipLaunchKernelGGL(FlxcorSendKernel, grid, dim3(64), 0, s, fc.dev, flux, fc.store, G, nb,
                     fc.ncomp);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status FlxcorSet(const Topology &t, const FluxCorrection &fc, double *flux, hipStream_t s) {
  if (fc.host.empty()) return PTH_COMPLETE;
  ProfScope prof("flxcor_set", s);
  const Geo G = MakeGeo(t, fc.ncomp);
  const int nb = static_cast<int>(t.blocks.size());
  const dim3 grid(static_cast<unsigned>(fc.host.size()), (fc.max_cells + 63) / 64);
  hipLaunchKernelGGL(FlxcorSetKernel, grid, dim3(64), 0, s, fc.dev, fc.store, flux, G, nb,
                     fc.ncomp);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

bool ReconFluxSupported(const Topology &t, int V) {
  const Extent &f = t.fine;
  const bool cubic = t.ndim == 3 && f.n[0] == f.n[1] && f.n[1] == f.n[2] &&
                     f.g[0] == f.g[1] && f.g[1] == f.g[2];
  return cubic && (f.n[0] == 16 || f.n[0] == 8) && f.g[0] == 4 && V >= 4 && V <= 13;
}

pth_status BurgersReconFlux(const Topology &t, int V, const double *U, double *flux, int recon,
                            int part, int ppw, hipStream_t s) {
  if (!ReconFluxSupported(t, V) || part < 0 || part > 2 || (ppw != 1 && ppw != 2 && ppw != 4)) {
    SetError("plane-tiled reconstruction: unsupported block shape, component count or split");
    return PTH_FAIL;
  }
  ProfScope prof(part == 0 ? "recon_flux" : part == 1 ? "recon_flux_inner" : "recon_flux_outer",
                 s);
  const int nb = static_cast<int>(t.blocks.size());
  const int N = t.fine.n[0];
  const dim3 grid(static_cast<unsigned>(3 * (N / ppw) * nb));
#define PTH_RF_LAUNCH(NN, NV, PT, PW)                                                         \
  hipLaunchKernelGGL((ReconFluxKernel<NN, 4, NV, PT, PW>), grid, dim3(kWG), 0, s, U, flux, nb, \
                     recon)
#define PTH_RF_PPW(NN, NV, PT)                                                                \
  if (ppw == 1)                                                                               \
    PTH_RF_LAUNCH(NN, NV, PT, 1);                                                             \
  else if (ppw == 2)                                                                          \
    PTH_RF_LAUNCH(NN, NV, PT, 2);                                                             \
  else                                                                                        \
    PTH_RF_LAUNCH(NN, NV, PT, 4);
#define PTH_RF_PARTS(NN, NV)                                                                  \
  if (part == 0) {                                                                            \
    PTH_RF_PPW(NN, NV, 0)                                                                     \
  } else if (part == 1) {                                                                     \
    PTH_RF_PPW(NN, NV, 1)                                                                     \
  } else {                                                                                    \
    PTH_RF_PPW(NN, NV, 2)                                                                     \
  }
#define PTH_RF_CASE(NV)                                                                       \
  case NV:                                                                                    \
    if (N == 16) {                                                                            \
      PTH_RF_PARTS(16, NV)                                                                    \
    } else {                                                                                  \
      PTH_RF_PARTS(8, NV)                                                                     \
    }                                                                                         \
    break;
  switch (V) {
    PTH_RF_CASE(4)
    PTH_RF_CASE(5)
    PTH_RF_CASE(6)
    PTH_RF_CASE(7)
    PTH_RF_CASE(8)
    PTH_RF_CASE(9)
    PTH_RF_CASE(10)
    PTH_RF_CASE(11)
    PTH_RF_CASE(12)
    PTH_RF_CASE(13)
  }
#undef PTH_RF_CASE
#undef PTH_RF_PARTS
#undef PTH_RF_PPW
#undef PTH_RF_LAUNCH
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

bool BurgersSweepSupported(const Topology &t, int V) { return ReconFluxSupported(t, V); }

// Workgroups of a kernel resident on the whole device at once (occupancy x compute units).
int ResidentGroups(const void *kernel, int threads) {
  static std::map<const void *, int> cache;
  auto it = cache.find(kernel);
  if (it != cache.end()) return it->second;
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess ||
      per < 1 || cus < 1)
    return -1;
  cache[kernel] = per * cus;
  return per * cus;
}

pth_status BurgersSweep(const Topology &t, int V, const double *U, const double *base,
                        double *out, double *derived, double *partial, int max_partials,
                        const BlockCoefs *coef, double beta, const double *dtvar, int recon,
                        const FluxCorrection *fc, int waves, int spread, const int *push,
                        int *groups, hipStream_t s) {
  if (!BurgersSweepSupported(t, V) || (waves != 3 && waves != 4) || spread < 1 || spread > 3) {
    SetError("block sweep: unsupported block shape or component count");
    return PTH_FAIL;
  }
  ProfScope prof("block_sweep", s);
  const int nb = static_cast<int>(t.blocks.size());
  const int N = t.fine.n[0];
  const int *fover = fc ? fc->dev_face : nullptr;
  const double *fcor = fc ? fc->store : nullptr;
  int grid = 0;
#define PTH_SW_LAUNCH_W(NN, NV, W)                                                            \
  {                                                                                           \
    const int r = ResidentGroups(reinterpret_cast<const void *>(&BlockSweepKernel<NN, 4, NV, W>), \
                                 NN * NN);                                                    \
    if (r < 1) {                                                                              \
      SetError("block sweep: occupancy query failed");                                        \
      return PTH_FAIL;                                                                        \
    }                                                                                         \
    grid = std::min<std::int64_t>(std::min(std::max(r / spread, 1), max_partials),             \
                                  static_cast<std::int64_t>(nb) * NN);                        \
    hipLaunchKernelGGL((BlockSweepKernel<NN, 4, NV, W>), dim3(grid), dim3(NN * NN), 0, s, U,  \
                       base, out, derived, partial, coef, nb, beta, dtvar, recon, fover, fcor, \
                       push);                                                                 \
  }
#define PTH_SW_LAUNCH(NN, NV)                                                                 \
  if (waves == 4)                                                                             \
    PTH_SW_LAUNCH_W(NN, NV, (NV <= 8 ? 4 : 2))                                                \
  else                                                                                        \
    PTH_SW_LAUNCH_W(NN, NV, (NV <= 8 ? 3 : 2))
#define PTH_SW_CASE(NV)                                                                       \
  case NV:                                                                                    \
    if (N == 16) {                                                                            \
      PTH_SW_LAUNCH(16, NV)                                                                   \
    } else {                                                                                  \
      PTH_SW_LAUNCH(8, NV)                                                                    \
    }                                                                                         \
    break;
  switch (V) {
    PTH_SW_CASE(4)
    PTH_SW_CASE(5)
    PTH_SW_CASE(6)
    PTH_SW_CASE(7)
    PTH_SW_CASE(8)
    PTH_SW_CASE(9)
    PTH_SW_CASE(10)
    PTH_SW_CASE(11)
    PTH_SW_CASE(12)
    PTH_SW_CASE(13)
  }
#undef PTH_SW_CASE
#undef PTH_SW_LAUNCH
#undef PTH_SW_LAUNCH_W
  PTH_HIP(hipGetLastError());
  if (groups) *groups = grid;
  return PTH_COMPLETE;
}

pth_status FlxcorDirect(const Topology &t, int V, const FluxCorrection &fc, const double *U,
                        int recon, hipStream_t s) {
  if (fc.host.empty()) return PTH_COMPLETE;
  if (V < 4 || V > 16) {
    SetError("flux correction supports 4 to 16 components");
    return PTH_FAIL;
  }
  ProfScope prof("flxcor_direct", s);
  const Geo G = MakeGeo(t, V);
  const dim3 grid(static_cast<unsigned>(fc.host.size()), (4 * fc.max_cells + 63) / 64);
#define PTH_FD_CASE(NV)                                                                       \
  case NV:                                                                                    \
    hipLaunchKernelGGL(FlxcorDirectKernel<NV>, grid, dim3(64), 0, s, fc.dev, U, fc.store, G,  \
                       recon);                                                                \
    break;
  switch (V) {
    PTH_FD_CASE(4)
    PTH_FD_CASE(5)
    PTH_FD_CASE(6)
    PTH_FD_CASE(7)
    PTH_FD_CASE(8)
    PTH_FD_CASE(9)
    PTH_FD_CASE(10)
    PTH_FD_CASE(11)
    PTH_FD_CASE(12)
    PTH_FD_CASE(13)
    PTH_FD_CASE(14)
    PTH_FD_CASE(15)
    PTH_FD_CASE(16)
  }
#undef PTH_FD_CASE
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersFluxUpdate(const Topology &t, int V, const double *flux, const double *U0,
                             const double *base, double *out, double *derived, double *partial,
                             const BlockCoefs *coef, double beta, const double *dtvar,
                             hipStream_t s) {
  if (V > kMaxV || V < 4) {
    SetError("flux update supports 4 to 16 components");
    return PTH_FAIL;
  }
  ProfScope prof("flux_update", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const int cells = t.fine.n[0] * t.fine.n[1] * t.fine.n[2];
  const dim3 grid(((cells + kWG - 1) / kWG) * nb);
#define PTH_FU_CASE(NV)                                                                       \
  case NV:                                                                                    \
    hipLaunchKernelGGL(FluxUpdateKernel<NV>, grid, dim3(kWG), 0, s, flux, U0, base, out,      \
                       derived, partial, coef, G, nb, beta, dtvar);                           \
    break;
  switch (V) {
    PTH_FU_CASE(4)
    PTH_FU_CASE(5)
    PTH_FU_CASE(6)
    PTH_FU_CASE(7)
    PTH_FU_CASE(8)
    PTH_FU_CASE(9)
    PTH_FU_CASE(10)
    PTH_FU_CASE(11)
    PTH_FU_CASE(12)
    PTH_FU_CASE(13)
    PTH_FU_CASE(14)
    PTH_FU_CASE(15)
    PTH_FU_CASE(16)
  }
#undef PTH_FU_CASE
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersMassHistory(const Topology &t, int V, const double *U,
                              const DeviceBlockGeom *geom, const BlockCoefs *coef,
                              const double mesh_lo[3], const double mesh_hi[3], double *scratch,
                              int nscratch, double *out, hipStream_t s) {
  ProfScope prof("mass_history", s);
  // octants in the reference order (burgers_package.cpp:104-119): x outer, then y, then z
  Octants O;
  double mid[3], vol = 1.0;
  for (int d = 0; d < 3; ++d) {
    vol *= (mesh_hi[d] - mesh_lo[d]);
    mid[d] = 0.5 * (mesh_lo[d] + mesh_hi[d]);
  }
  int o = 0;
  for (int s1 = 0; s1 < 2; ++s1)
    for (int s2 = 0; s2 < 2; ++s2)
      for (int s3 = 0; s3 < 2; ++s3, ++o) {
        const int side[3] = {s1, s2, s3};
        for (int d = 0; d < 3; ++d) {
          O.lo[o][d] = side[d] ? mid[d] : mesh_lo[d];
          O.hi[o][d] = side[d] ? mesh_hi[d] : mid[d];
        }
      }
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const std::int64_t total =
      static_cast<std::int64_t>(t.fine.n[0]) * t.fine.n[1] * t.fine.n[2] * V * nb;
  int grid = Grid(total);
  if (grid > nscratch / 8) grid = nscratch / 8;
  hipLaunchKernelGGL(MassHistoryKernel, dim3(grid), dim3(kWG), 0, s, U, geom, coef, G, nb, O,
                     vol, scratch);
  hipLaunchKernelGGL(SumPartials8Kernel, dim3(1), dim3(64), 0, s, scratch, grid, out);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersClock(const double *partial, int n, double cfl, double *dtvar, int advance,
                        hipStream_t s) {
  ProfScope prof("clock", s);
  hipLaunchKernelGGL(ClockKernel, dim3(1), dim3(kClockWG), 0, s, partial, n, cfl, dtvar,
                     advance);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status BurgersDtMin(const Topology &t, int V, const double *U, const BlockCoefs *coef,
                        double *partial, int npartial, double *out, hipStream_t s) {
  ProfScope prof("dt_reduce", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const std::int64_t total =
      static_cast<std::int64_t>(t.fine.n[0]) * t.fine.n[1] * t.fine.n[2] * nb;
  int grid = Grid(total);
  if (grid > npartial) grid = npartial;
  hipLaunchKernelGGL(DtKernel, dim3(grid), dim3(kWG), 0, s, U, coef, G, nb, partial);
  hipLaunchKernelGGL(MinFinalKernel, dim3(1), dim3(64), 0, s, partial, grid, out);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

}  // namespace pth
