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
// Expression order matches the reference operator by operator and this file is built with
// -ffp-contract=off, so results are reproducible bit for bit by an IEEE CPU evaluation (the
// block-sweep form, in pth_sweep.hip, is built with contraction instead).
// Arrays are [block][comp][k][j][i] bricks of (n+2g)^3 with i fastest; consecutive lanes
// take consecutive i, so every load and store is a coalesced run along x.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "pth_burgers.hpp"
#include "pth_device.hpp"
#include "pth_stencil.hpp"

namespace pth {

namespace {

constexpr int kWG = 256;

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

// cfl times the minimum of the partials, into out[0] (the product the host forms after
// MinFinalKernel in the synchronous form, so the two give the same bits)
__global__ void __launch_bounds__(64) MinScaleKernel(const double *__restrict__ partial, int n,
                                                     double cfl, double *__restrict__ out) {
  double m = DBL_MAX;
  for (int t = threadIdx.x; t < n; t += 64) m = DMin(m, partial[t]);
  m = WaveMin(m);
  if (threadIdx.x == 0) out[0] = cfl * m;
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

// Final CFL minimum and EvolutionDriver::SetGlobalTimeStep on device (driver.cpp:210-269,
// dt_factor 2), with SimTime::KeepGoing's stop at tlim (basic_types.hpp:241).
// dtvar = {dt of the next cycle, time, estimate, dt of the last cycle, tm.dt}.
// advance 0 (start, ncycle 0): tm.dt is still the initial maximum and dt_init applies;
// advance 1 (end of a cycle): time += dt and the cycle's dt is kept before the rule;
// advance 2 (after a remesh): the estimate of the new blocks replaces the old one and the rule
// starts again from the dt of the last cycle.
// The loop of EvolutionDriver::Execute ends once time >= tlim. A cycle queued after that runs
// with dt = 0 (slot 0), which leaves the state, the time and the clock unchanged, so a host
// loop may queue cycles without reading the clock; slot 4 keeps the reference's tm.dt.
// One workgroup of kClockWG threads: strided partial minima, wavefront shuffles, then LDS.
constexpr int kClockWG = 1024;
__global__ void __launch_bounds__(kClockWG)
    ClockKernel(const double *__restrict__ partial, int n, double cfl, double *dtvar,
                int advance, ClockLimits L) {
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
    double time = dtvar[1];
    if (advance == 1) {
      if (!(time < L.tlim)) return;  // a cycle after the end: nothing moved
      time = time + dtvar[0];
      dtvar[1] = time;
      dtvar[3] = dtvar[0];
    }
    double dt = advance == 0 ? dtvar[4] : dtvar[3];
    if (L.dt_force > 0.0) {
      dt = L.dt_force;
    } else if (advance == 0 && L.dt_init_force && L.dt_init > 0.0) {
      dt = L.dt_init;
    } else {
      if (dt < 0.1 * DBL_MAX) dt *= 2.0;
      if (advance == 0) dt = DMin(dt, L.dt_init);
      dt = DMin(dt, est);
      dt = DMin(dt, L.dt_user);
      dt = DMax(L.dt_floor, DMin(dt, L.dt_ceil));
    }
    if (time < L.tlim && (L.tlim - time) < dt) dt = L.tlim - time;
    dtvar[4] = dt;
    dtvar[0] = time < L.tlim ? dt : 0.0;
    dtvar[2] = est;
  }
}

int Grid(std::int64_t total) {
  const std::int64_t want = (total + kWG - 1) / kWG;
  return static_cast<int>(want < 8192 ? want : 8192);
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
                        const ClockLimits &lim, hipStream_t s) {
  ProfScope prof("clock", s);
  hipLaunchKernelGGL(ClockKernel, dim3(1), dim3(kClockWG), 0, s, partial, n, cfl, dtvar,
                     advance, lim);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

int BurgersDtPartials(const Topology &t) {
  return Grid(static_cast<std::int64_t>(t.fine.n[0]) * t.fine.n[1] * t.fine.n[2] *
              static_cast<std::int64_t>(t.blocks.size()));
}

pth_status BurgersDtToDevice(const Topology &t, int V, const double *U, const BlockCoefs *coef,
                             double *partial, int npartial, double cfl, double *out,
                             hipStream_t s) {
  ProfScope prof("dt_reduce", s);
  const Geo G = MakeGeo(t, V);
  const int nb = static_cast<int>(t.blocks.size());
  const int grid = BurgersDtPartials(t);
  if (grid > npartial) {
    SetError("CFL reduction: more partials than the scratch holds");
    return PTH_FAIL;
  }
  hipLaunchKernelGGL(DtKernel, dim3(grid), dim3(kWG), 0, s, U, coef, G, nb, partial);
  hipLaunchKernelGGL(MinScaleKernel, dim3(1), dim3(64), 0, s, partial, grid, cfl, out);
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
