// Block-sweep form of the Burgers stage for gfx950, FP64: one launch per stage does the
// reconstruction, HLL fluxes, flux divergence, RK update, derived field and CFL minimum
// (benchmarks/burgers/burgers_package.cpp:143-200, 236-401; src/interface/update.cpp:63-84,
// update.hpp:43-90), plus the direct flux-correction send of the same stage.
//
// This file is built with FMA contraction (-ffp-contract=fast), as a CUDA build of the
// reference is by default (--fmad=true): a multiply feeding an add becomes one fused
// instruction, which takes a quarter of the FP64 instructions out of this issue-bound kernel.
// Its results therefore agree with an IEEE CPU evaluation to rounding (about 1e-15 relative
// after several cycles) rather than bit for bit; the task and plane-tiled forms in
// pth_burgers.hip are built without contraction and stay bit-exact. The flux-correction send
// of this form is here too, so the fine fluxes it restricts are the same bits as the ones the
// sweep applies to the fine blocks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <map>
#include <type_traits>

#include "pth_burgers.hpp"
#include "pth_device.hpp"
#include "pth_stencil.hpp"

namespace pth {

namespace {

// HllGen with the scalar components multiplied by the reciprocal the velocity components
// already use (burgers_package.cpp:307-401 divides them by the same denominator).
template <int V, class LS, class RS>
__device__ __forceinline__ void SweepHll(LS l, RS r, int d, double *F) {
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
    F[n] = (sr * upl * ql - sl * upr * qr + sl * sr * (qr - ql)) * islsr;
  }
}

// ---- helpers of the 16^3 block sweep ---------------------------------------------------
//
// Reciprocal by v_rcp_f64 and two Newton-Raphson steps r <- r + r (1 - x r) (each step
// roughly doubles the correct bits; the hardware estimate is good to well over 20 bits, so
// the result is within about an ulp), five instructions instead of the ten of an IEEE divide.
// The callers' arguments are positive normal numbers (sums of squares plus robust::EPS, or
// HLL denominators that are >= 1 whenever the speeds' product underflows).
__device__ __forceinline__ double Recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// WENO5Z of recon.hpp:28-99 for the 16^3 block sweep, straight-line code with six
// reciprocals by Recip instead of eight divides: one reciprocal of be0 be1 be2 gives the three
// 1 / be_k; each side's weights enter through 1 / (S D) (S = w0 + w1 + w2, D = g2 w0 w1 +
// g1 w0 w2 + g0 w1 w2, so that ql = N D / (S D) and al = 3 w0 w1 w2 / (S D) + eps); alin's
// harmonic mean takes the sixth. The values are the reference's to rounding. Range: be_k >=
// eps and w_k <= g_k (1 + tau5 / eps), so be0 be1 be2 stays finite below a jump in q of about
// 1e50 and S D ~ w^3 below a jump of about 1e43 beside a flat stencil; the reference's own
// expression overflows (w^2) near 1e69.
__device__ __forceinline__ void Weno5ZFast(double q0, double q1, double q2, double q3,
                                           double q4, double &ql, double &qr) {
  const double a00 = 1.0 / 3.0, a01 = -7.0 / 6.0, a02 = 11.0 / 6.0;
  const double a10 = -1.0 / 6.0, a11 = 5.0 / 6.0, a12 = 1.0 / 3.0;
  const double a20 = 1.0 / 3.0, a21 = 5.0 / 6.0, a22 = -1.0 / 6.0;
  const double g0 = 0.1, g1 = 0.6, g2 = 0.3;
  const double eps = 10.0 * DBL_EPSILON;  // robust::EPS (src/utils/robust.hpp:38-41)
  const double t13 = 13.0 / 3.0;
  double a = q0 - 2 * q1 + q2;
  double b = q0 - 4.0 * q1 + 3.0 * q2;
  const double be0 = t13 * a * a + b * b + eps;
  a = q1 - 2.0 * q2 + q3;
  b = q3 - q1;
  const double be1 = t13 * a * a + b * b + eps;
  a = q2 - 2.0 * q3 + q4;
  b = q4 - 4.0 * q3 + 3.0 * q2;
  const double be2 = t13 * a * a + b * b + eps;
  const double tau5 = fabs(be2 - be0);
  const double p12 = be1 * be2, p02 = be0 * be2, p01 = be0 * be1;
  const double rp = Recip(be0 * p12);
  const double z0 = (be0 + tau5) * (rp * p12);
  const double z1 = (be1 + tau5) * (rp * p02);
  const double z2 = (be2 + tau5) * (rp * p01);
  const double l0 = g0 * z0 + eps, l1 = g1 * z1 + eps, l2 = g2 * z2 + eps;
  const double nl = l0 * (a00 * q0 + a01 * q1 + a02 * q2) + l1 * (a10 * q1 + a11 * q2 + a12 * q3) +
                    l2 * (a20 * q2 + a21 * q3 + a22 * q4);
  const double dl = g2 * l0 * l1 + g1 * l0 * l2 + g0 * l1 * l2;
  const double il = Recip((l0 + l1 + l2) * dl);  // 1 / (S_l D_l)
  const double r0 = g0 * z2 + eps, r1 = g1 * z1 + eps, r2 = g2 * z0 + eps;
  const double nr = r0 * (a00 * q4 + a01 * q3 + a02 * q2) + r1 * (a10 * q3 + a11 * q2 + a12 * q1) +
                    r2 * (a20 * q2 + a21 * q1 + a22 * q0);
  const double dr = g2 * r0 * r1 + g1 * r0 * r2 + g0 * r1 * r2;
  const double ir = Recip((r0 + r1 + r2) * dr);  // 1 / (S_r D_r)
  const double wl = nl * dl * il, wr = nr * dr * ir;
  const double al = 3.0 * l0 * l1 * l2 * il + eps;
  const double ar = 3.0 * r0 * r1 * r2 * ir + eps;
  double dq = q3 - q2;
  dq = 0.5 * Mc(q2 - q1, dq);
  const double alin = 2.0 * al * ar * Recip(al + ar);
  ql = alin * wl + (1.0 - alin) * (q2 + dq);
  qr = alin * wr + (1.0 - alin) * (q2 - dq);
}

__device__ __forceinline__ void FastStates(int recon, double q0, double q1, double q2, double q3,
                                           double q4, double &ql, double &qr) {
  if (recon == 0)
    Weno5ZFast(q0, q1, q2, q3, q4, ql, qr);
  else
    LinearRecon(q1, q2, q3, ql, qr);
}

// CellStates with the sweep's WENO5Z (the 8^3 block sweep and its flux-correction send).
__device__ __forceinline__ void SweepStates(const double *q, unsigned c, unsigned st, int recon,
                                            double &ql, double &qr) {
  if (recon == 0)
    Weno5ZFast(LoadAt(q, c - 2 * st), LoadAt(q, c - st), LoadAt(q, c), LoadAt(q, c + st),
               LoadAt(q, c + 2 * st), ql, qr);
  else
    LinearRecon(LoadAt(q, c - st), LoadAt(q, c), LoadAt(q, c + st), ql, qr);
}

// HLL of burgers_package.cpp:307-401 (lr_to_flux, burgers_package.hpp:33-43) as two
// coefficients per face: with sl, sr the signal speeds and islsr = 1 / (sr - sl + (sl sr ==
// 0)), the scalar flux (sr upl ql - sl upr qr + sl sr (qr - ql)) islsr equals a ql + b qr with
// a = islsr sr (upl - sl), b = islsr sl (sr - upr), and the velocity flux is half of the same
// combination of ul, ur. So the coupling of the components goes through (a, b) only.
__device__ __forceinline__ void HllCoef(double upl, double upr, double &a, double &b) {
  const double sl = DMin(DMin(upl, upr), 0.0);
  const double sr = DMax(DMax(upl, upr), 0.0);
  const double is = Recip(sr - sl + (sl * sr == 0.0));
  a = (is * sr) * (upl - sl);
  b = (is * sl) * (sr - upr);
}

// flux of component n from the face coefficients and its left / right states
__device__ __forceinline__ double HllFluxN(int n, double a, double b, double l, double r) {
  if (n < 3) return fma(0.5 * a, l, (0.5 * b) * r);
  return fma(a, l, b * r);
}

__device__ __forceinline__ int DMinI(int a, int b) { return b < a ? b : a; }

// value of lane - 1 / lane + 1 within each 16-lane DPP row (the row's first / last lane
// gets 0; callers replace those)
__device__ __forceinline__ double FromLeft16(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x111, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x111, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double FromRight16(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x101, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x101, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// flxcor_send of the block-sweep stage: four adjacent threads per coarse face cell of a patch,
// one per fine face (p, q), evaluate the HLL flux of their face straight from the fine block's
// state (the left state of face f is ql of cell f - 1, the right state qr of cell f, as
// CalculateFluxes stores them); the first of the four gathers the other three by lane
// shuffles and writes their restriction to the correction store.
template <int V, bool RING>
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
    if (RING) {
      // the expressions of BlockSweep16Kernel: FastStates and the face coefficients
#pragma unroll
      for (int n = 0; n < V; ++n) {
        const double *q = Ub + n * cs;
        double x, y;
        FastStates(recon, LoadAt(q, o - 3 * st), LoadAt(q, o - 2 * st), LoadAt(q, o - st),
                   LoadAt(q, o), LoadAt(q, o + st), l[n], x);
        FastStates(recon, LoadAt(q, o - 2 * st), LoadAt(q, o - st), LoadAt(q, o),
                   LoadAt(q, o + st), LoadAt(q, o + 2 * st), y, r[n]);
      }
      double a, bb;
      HllCoef(l[P.d], r[P.d], a, bb);
#pragma unroll
      for (int n = 0; n < V; ++n) F[n] = HllFluxN(n, a, bb, l[n], r[n]);
    } else {
#pragma unroll
      for (int n = 0; n < V; ++n) {
        double x, y;
        SweepStates(Ub + n * cs, o - st, st, recon, l[n], x);
        SweepStates(Ub + n * cs, o, st, recon, y, r[n]);
      }
      SweepHll<V>([&](int n) { return l[n]; }, [&](int n) { return r[n]; }, P.d, F);
    }
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
// the other forms up to FMA contraction and the folded divides, so the stage output agrees
// with theirs to rounding, and no flux or face state reaches HBM. On
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
        SweepStates(Ub + n * cs, c, sz, recon, sl[n * P + u], sr[n * P + u]);
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
    SweepHll<V>([&](int n) { return zl[n]; }, [&](int n) { return sr[n * P + threadIdx.x]; }, 2, zf);
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
        SweepStates(Ub + n * cs, c, sx, recon, sl[t], sr[t]);
      }
      __syncthreads();
      {
        // every x face once: each thread the high face of its cell, threads 0 .. N - 1 also
        // the low face of row threadIdx.x; fluxes to LDS as [n][row][face]
        double F[V];
        const double *L = sl + ty * R + tx, *Rr = sr + ty * R + tx;
        SweepHll<V>([&](int n) { return L[n * P + 1]; }, [&](int n) { return Rr[n * P + 2]; }, 0,
                  F);
        if (tx == N - 1 && ox1 >= 0)
#pragma unroll
          for (int n = 0; n < V; ++n) F[n] = fcor[ox1 + ty + k * N + n * comp];
#pragma unroll
        for (int n = 0; n < V; ++n) fx[n * FP + ty * (N + 1) + tx + 1] = F[n];
        if (static_cast<int>(threadIdx.x) < N) {
          const int row = threadIdx.x;
          SweepHll<V>([&](int n) { return sl[n * P + row * R]; },
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
        SweepStates(Ub + n * cs, c, sy, recon, sl[t], sr[t]);
      }
      __syncthreads();  // also: every thread has read its x fluxes
      {
        // every y face once, as [n][face][i]
        double F[V];
        const double *L = sl + ty * N + tx, *Rr = sr + ty * N + tx;
        SweepHll<V>([&](int n) { return L[n * P + N]; }, [&](int n) { return Rr[n * P + 2 * N]; },
                  1, F);
        if (ty == N - 1 && oy1 >= 0)
#pragma unroll
          for (int n = 0; n < V; ++n) F[n] = fcor[oy1 + tx + k * N + n * comp];
#pragma unroll
        for (int n = 0; n < V; ++n) fx[n * FP + (ty + 1) * N + tx] = F[n];
        if (static_cast<int>(threadIdx.x) < N) {
          const int col = threadIdx.x;
          SweepHll<V>([&](int n) { return sl[n * P + col]; },
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
        SweepHll<V>([&](int n) { return zl[n]; }, [&](int n) { return sr[n * P + threadIdx.x]; }, 2,
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

// ---- block sweep for 16^3 blocks: register-ring z stencils, LDS plane, DPP neighbours ------
//
// One workgroup of 256 threads marches the z planes of a block; thread (tx, ty) owns column
// (i, j) = (tx, ty). Per plane kk:
//   * the plane's U (all components, with 3-cell x and y halos) is staged in LDS; its interior
//     comes from the thread's register ring, the halos from HBM;
//   * x: each thread reconstructs its own cell (stencil read from LDS); the neighbouring
//     cells' states come by DPP row shifts (a row of 16 cells is a 16-lane DPP row), those of
//     the cells beyond the row ends from a small pass over all components; the face
//     coefficients (a, b) of HllCoef follow from the x velocity's states, then every
//     component's two face fluxes are two multiply-adds each;
//   * y: the same with the thread mapping transposed (thread (tx, ty) takes cell i = ty,
//     j = tx, so a column is a DPP row); the y differences reach the owning thread through
//     LDS;
//   * z: each thread keeps U of its column at planes kk - 1, kk + 1, kk + 2, kk + 3 in
//     registers (kk itself is in LDS), so the z stencil of the cell above needs no load; the
//     high z face's flux becomes the next plane's low face flux; the load of plane kk + 4 is
//     issued a whole plane ahead of its use.
// The divergence adds x, y and z in the order of FluxDivHelper (update.hpp:43-57); the RK
// averaging and update, derived field, CFL minimum, flux-correction faces and same-level push
// are those of BlockSweepKernel. Each U element is read from HBM once per workgroup (plus
// the halos and the segment's first planes), and face states and fluxes never leave the CU.
template <int V, int MINW, int RECON>
__global__ void __launch_bounds__(256, MINW)
    BlockSweep16Kernel(const double *__restrict__ U, const double *base, double *out,
                       double *__restrict__ derived, double *__restrict__ partial,
                       const BlockCoefs *__restrict__ coef, int nblocks, double beta,
                       const double *__restrict__ dtvar, const int *__restrict__ fover,
                       const double *__restrict__ fcor, const int *__restrict__ push,
                       const int *__restrict__ list) {
  constexpr int N = 16, NG = 4, E = N + 2 * NG, NT = N * N;
  constexpr int H = 3;             // halo of the staged plane
  constexpr int PW = N + 2 * H;    // staged rows / columns
  constexpr int PP = PW + 1;       // row pitch (odd, fewer bank conflicts on column reads)
  constexpr int PS = PW * PP;      // one component of the staged plane
  constexpr int SP = N + 1;        // pitch of the difference tile
  constexpr int NX = V * N * 2;    // beyond-the-row states of one direction
  constexpr int NH = V * 12 * N;   // halo cells of one staged plane (all components)
  constexpr int QH = (NH + NT - 1) / NT;  // halo cells per thread
  constexpr std::int64_t cs = static_cast<std::int64_t>(E) * E * E;
  constexpr std::int64_t sz = static_cast<std::int64_t>(E) * E;
  constexpr int comp = N * N;      // correction-store stride between components
  __shared__ double pl[V * PS];      // [n][j + H][i + H]
  __shared__ double ext[2 * NX];     // [dir][n][line][side]: side 0 ql of cell -1, 1 qr of cell N
  __shared__ double sd[V * N * SP];  // x + y differences [n][j][i], then the new state
  __shared__ double zs[2 * V * NT];  // this thread's z state: [0][n] ql of its cell, [1][n] flux
                                     // of its low z face (own slots, no barriers)
  __shared__ double fxy[2 * V * 2 * N];  // corrected x / y face fluxes of the plane,
                                          // [dir][n][side][line] (blocks with such faces)
  __shared__ double wmin[NT / 64];
  const int tx = threadIdx.x % N, ty = threadIdx.x / N;
  const double bdt = beta * dtvar[0];
  const double w2 = 1.0 - beta;
  const bool same_base = base == U;
  double m = DBL_MAX;
  const std::int64_t T = static_cast<std::int64_t>(nblocks) * N;
  std::int64_t pnext = T * blockIdx.x / gridDim.x;
  const std::int64_t pend = T * (blockIdx.x + 1) / gridDim.x;
  const double *own_pl = pl + (ty + H) * PP + tx + H;  // own cell, component 0
  const int own_sd = ty * SP + tx;                      // own cell in the difference tile
  // halo cell q of this thread (q < QH): its component, offset within a plane of the brick
  // and place in the staged plane; cells past the last share its address (loads are
  // unconditional, so the compiler can count them in its memory waits)
  auto halo = [&](int q, std::int64_t &goff, int &loff) {
    const int t = DMinI(static_cast<int>(threadIdx.x) + q * NT, NH - 1);
    const int n = t / (12 * N), r = t % (12 * N);
    int i, j;
    if (r < 6 * N) {  // y strips: rows -3 .. -1 and N .. N + 2, all interior columns
      const int yy = r / N;
      i = r % N;
      j = yy < 3 ? yy - 3 : N + yy - 3;
    } else {          // x strips: all interior rows, columns -3 .. -1 and N .. N + 2
      const int qq = r - 6 * N;
      j = qq / 6;
      const int c = qq % 6;
      i = c < 3 ? c - 3 : N + c - 3;
    }
    goff = n * cs + static_cast<std::int64_t>(NG + j) * E + NG + i;
    loff = n * PS + (j + H) * PP + i + H;
  };
  // the new state of plane kk (in the thread's own slots of sd) to out, the same-level
  // neighbours' ghosts, the derived field and the CFL minimum
  auto flush = [&](int b, int k) {
    const int kk = NG + k;
    const std::int64_t cell = static_cast<std::int64_t>(kk) * sz + (NG + ty) * E + NG + tx;
    const std::int64_t c0 = static_cast<std::int64_t>(b) * V * cs + cell;
    double u1[V];
#pragma unroll
    for (int n = 0; n < V; ++n) {
      u1[n] = sd[n * N * SP + own_sd];
      out[c0 + n * cs] = u1[n];
    }
    if (push) {
      // same-level ghost copies: a cell within NG of a face is a ghost cell of the
      // same-level neighbours across that face, edge or corner (push[27 b + direction]), at
      // its index shifted by N along each crossed dimension
      const int lx = tx < NG ? -1 : (tx >= N - NG ? 1 : 0);
      const int ly = ty < NG ? -1 : (ty >= N - NG ? 1 : 0);
      const int lz = k < NG ? -1 : (k >= N - NG ? 1 : 0);
      for (int az_ = 0; az_ < (lz ? 2 : 1); ++az_)
        for (int ay = 0; ay < (ly ? 2 : 1); ++ay)
          for (int ax = 0; ax < (lx ? 2 : 1); ++ax) {
            const int dx = ax ? lx : 0, dy = ay ? ly : 0, dz = az_ ? lz : 0;
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
    const BlockCoefs &Cb = coef[b];
    double sp = fabs(u1[0]) / Cb.dx[0];
    sp = sp + fabs(u1[1]) / Cb.dx[1];
    sp = sp + fabs(u1[2]) / Cb.dx[2];
    m = DMin(m, 1.0 / sp);
  };
  while (pnext < pend) {
    // the planes of the launch run over its blocks in order: all blocks, or the listed ones
    const int b = list ? list[pnext / N] : static_cast<int>(pnext / N);
    const int k0 = static_cast<int>(pnext % N);
    const int k1 = static_cast<int>(pend - pnext < N - k0 ? k0 + (pend - pnext) : N);
    pnext += k1 - k0;
    const double *Ub = U + static_cast<std::int64_t>(b) * V * cs;
    const BlockCoefs &C = coef[b];
    const double ivol = 1.0 / C.vol;
    const int ox0 = fover ? fover[6 * b + 0] : -1, ox1 = fover ? fover[6 * b + 1] : -1;
    const int oy0 = fover ? fover[6 * b + 2] : -1, oy1 = fover ? fover[6 * b + 3] : -1;
    const int oz0 = fover ? fover[6 * b + 4] : -1, oz1 = fover ? fover[6 * b + 5] : -1;
    const std::int64_t col = static_cast<std::int64_t>(NG + ty) * E + NG + tx;  // in a plane
    const double *Uc = Ub + col;  // this thread's column, component 0, plane 0
    const int kk0 = NG + k0;
    // halo of the first plane
    double hv[QH];
#pragma unroll
    for (int q = 0; q < QH; ++q) {
      std::int64_t go;
      int lo;
      halo(q, go, lo);
      hv[q] = Ub[go + kk0 * sz];
    }
    // ---- z state at the segment's first plane kk0: ql of cell kk0 (zl) and the flux of face
    // kk0 - 1/2 (flo); the z velocity first gives the face coefficients. The per-component
    // part runs as a loop through LDS (once per segment, kept out of the unrolled code).
    double az, bz;
    {
      const double *q = Uc + 2 * cs + static_cast<std::int64_t>(kk0 - 3) * sz;
      double qm, qx, ql, qr;
      FastStates(RECON, q[0], q[sz], q[2 * sz], q[3 * sz], q[4 * sz], qm, qx);
      FastStates(RECON, q[sz], q[2 * sz], q[3 * sz], q[4 * sz], q[5 * sz], ql, qr);
      HllCoef(qm, qr, az, bz);
    }
#pragma unroll 1
    for (int n = 0; n < V; ++n) {
      const double *q = Uc + n * cs + static_cast<std::int64_t>(kk0 - 3) * sz;
      double qm, qx, ql, qr;
      FastStates(RECON, q[0], q[sz], q[2 * sz], q[3 * sz], q[4 * sz], qm, qx);
      FastStates(RECON, q[sz], q[2 * sz], q[3 * sz], q[4 * sz], q[5 * sz], ql, qr);
      double f = HllFluxN(n, az, bz, qm, qr);
      if (k0 == 0 && oz0 >= 0) f = fcor[oz0 + tx + ty * N + n * comp];
      zs[(V + n) * NT + threadIdx.x] = f;
      zs[n * NT + threadIdx.x] = ql;
    }
    double r0[V], r1[V], r2[V], r3[V], r4[V];
#pragma unroll
    for (int n = 0; n < V; ++n) {
      const double *q = Uc + n * cs + static_cast<std::int64_t>(kk0 - 1) * sz;
      r0[n] = q[0];
      r1[n] = q[sz];
      r2[n] = q[2 * sz];
      r3[n] = q[3 * sz];
      r4[n] = q[4 * sz];
    }
    for (int k = k0; k < k1; ++k) {
      const int kk = NG + k;
      __syncthreads();  // the previous plane's (or the segment start's) LDS has been read
      // ---- stage plane kk: interior from the ring, halos from the prefetch
#pragma unroll
      for (int n = 0; n < V; ++n) pl[n * PS + (ty + H) * PP + tx + H] = r1[n];
#pragma unroll
      for (int q = 0; q < QH; ++q) {
        std::int64_t go;
        int lo;
        halo(q, go, lo);
        pl[lo] = hv[q];
      }
      // corrected fluxes of this plane's x and y faces (only blocks next to finer ones):
      // staged here, so the loops below read LDS and issue no conditional load
      if (ox0 >= 0 || ox1 >= 0 || oy0 >= 0 || oy1 >= 0) {
        for (int t = threadIdx.x; t < 2 * V * 2 * N; t += NT) {
          const int dir = t / (2 * V * N), r = t % (2 * V * N);
          const int n = r / (2 * N), side = (r / N) & 1, line = r % N;
          const int o = dir == 0 ? (side ? ox1 : ox0) : (side ? oy1 : oy0);
          fxy[t] = o >= 0 ? fcor[o + line + k * N + n * comp] : 0.0;
        }
      }
      // the previous plane's results leave now, well before the next wait on a load; then the
      // ring's top plane (kk + 3) for this plane's z stencils, issued after those stores so
      // that the wait for it counts only loads
      if (k > k0) {
        flush(b, k - 1);
#pragma unroll
        for (int n = 0; n < V; ++n) r4[n] = Uc[n * cs + static_cast<std::int64_t>(kk + 3) * sz];
      }
      __syncthreads();
      // ---- states of the cells just beyond each row (x) and column (y), every component
      for (int t = threadIdx.x; t < 2 * NX; t += NT) {
        const int dir = t / NX, r = t % NX;
        const int n = r / (2 * N), q = r % (2 * N);
        const int line = q >> 1, side = q & 1;
        const int cell = side ? N : -1;
        const int st = dir ? PP : 1;
        const double *c = pl + n * PS + (dir ? (cell + H) * PP + line + H
                                             : (line + H) * PP + cell + H);
        double ql, qr;
        FastStates(RECON, c[-2 * st], c[-st], c[0], c[st], c[2 * st], ql, qr);
        ext[t] = side ? qr : ql;
      }
      __syncthreads();
      // ---- x: own cell (tx, ty); differences to sd[n][ty][tx]
      {
        const double A = C.area[0];
        double ql0, qr0;
        FastStates(RECON, own_pl[-2], own_pl[-1], own_pl[0], own_pl[1], own_pl[2], ql0, qr0);
        double qlm = FromLeft16(ql0), qrp = FromRight16(qr0);
        if (tx == 0) qlm = ext[ty * 2];
        if (tx == N - 1) qrp = ext[ty * 2 + 1];
        double alo, blo, ahi, bhi;
        HllCoef(qlm, qr0, alo, blo);
        HllCoef(ql0, qrp, ahi, bhi);
#pragma unroll 2
        for (int n = 0; n < V; ++n) {
          double ql = ql0, qr = qr0;
          if (n > 0) {
            const double *cn = own_pl + n * PS;
            FastStates(RECON, cn[-2], cn[-1], cn[0], cn[1], cn[2], ql, qr);
          }
          double lm = FromLeft16(ql), rp = FromRight16(qr);
          if (tx == 0) lm = ext[(n * N + ty) * 2];
          if (tx == N - 1) rp = ext[(n * N + ty) * 2 + 1];
          double fl = HllFluxN(n, alo, blo, lm, qr);
          double fh = HllFluxN(n, ahi, bhi, ql, rp);
          if (tx == 0 && ox0 >= 0) fl = fxy[(n * 2 + 0) * N + ty];
          if (tx == N - 1 && ox1 >= 0) fh = fxy[(n * 2 + 1) * N + ty];
          sd[n * N * SP + own_sd] = A * fh - A * fl;
        }
      }
      __syncthreads();  // the x differences are in the tile
      // ---- y: cell i = ty, j = tx (a column per DPP row); adds to sd[n][j][i]
      {
        const double A = C.area[1];
        const double *c = pl + (tx + H) * PP + ty + H;
        double ql1, qr1;
        FastStates(RECON, c[PS - 2 * PP], c[PS - PP], c[PS], c[PS + PP], c[PS + 2 * PP], ql1,
                   qr1);
        double qlm = FromLeft16(ql1), qrp = FromRight16(qr1);
        if (tx == 0) qlm = ext[NX + (N + ty) * 2];
        if (tx == N - 1) qrp = ext[NX + (N + ty) * 2 + 1];
        double alo, blo, ahi, bhi;
        HllCoef(qlm, qr1, alo, blo);
        HllCoef(ql1, qrp, ahi, bhi);
#pragma unroll 2
        for (int n = 0; n < V; ++n) {
          double ql = ql1, qr = qr1;
          if (n != 1) {
            const double *cn = c + n * PS;
            FastStates(RECON, cn[-2 * PP], cn[-PP], cn[0], cn[PP], cn[2 * PP], ql, qr);
          }
          double lm = FromLeft16(ql), rp = FromRight16(qr);
          if (tx == 0) lm = ext[NX + (n * N + ty) * 2];
          if (tx == N - 1) rp = ext[NX + (n * N + ty) * 2 + 1];
          double fl = HllFluxN(n, alo, blo, lm, qr);
          double fh = HllFluxN(n, ahi, bhi, ql, rp);
          if (tx == 0 && oy0 >= 0) fl = fxy[((V + n) * 2 + 0) * N + ty];
          if (tx == N - 1 && oy1 >= 0) fh = fxy[((V + n) * 2 + 1) * N + ty];
          double &slot = sd[n * N * SP + tx * SP + ty];
          slot = slot + (A * fh - A * fl);
        }
      }
      __syncthreads();  // x + y differences complete
      // the base state of this plane (second stage), then the next plane's halo; both are
      // issued before this plane's results are stored, so no wait drains a store
      const std::int64_t c0 = static_cast<std::int64_t>(b) * V * cs +
                              static_cast<std::int64_t>(kk) * sz + col;
      double bv[V];
#pragma unroll
      for (int n = 0; n < V; ++n) bv[n] = same_base ? 0.0 : base[c0 + n * cs];
      if (k + 1 < k1) {
#pragma unroll
        for (int q = 0; q < QH; ++q) {
          std::int64_t go;
          int lo;
          halo(q, go, lo);
          hv[q] = Ub[go + (kk + 1) * sz];
        }
      }
      // ---- z: the cell above (kk + 1) from the ring; its low face closes this cell
      const double A = C.area[2];
      {
        double ql2, qr2;
        FastStates(RECON, r0[2], own_pl[2 * PS], r2[2], r3[2], r4[2], ql2, qr2);
        HllCoef(zs[2 * NT + threadIdx.x], qr2, az, bz);
      }
      // the z loop in two copies: the top plane of a block whose upper z face takes the
      // correction store's fluxes loads them; no other plane has a load in this loop
      auto zloop = [&](auto top_tag) {
        constexpr bool TOP = decltype(top_tag)::value;
  #pragma unroll
        for (int n = 0; n < V; ++n) {
          const double uo = own_pl[n * PS];
          double ql, qr;
          FastStates(RECON, r0[n], uo, r2[n], r3[n], r4[n], ql, qr);
          double &zl = zs[n * NT + threadIdx.x], &flo = zs[(V + n) * NT + threadIdx.x];
          double fh = HllFluxN(n, az, bz, zl, qr);
          if (TOP) fh = fcor[oz1 + tx + ty * N + n * comp];
          double &slot = sd[n * N * SP + own_sd];
          const double du = slot + (A * fh - A * flo);
          flo = fh;
          zl = ql;
          const double dudt = -du * ivol;
          const double avg = beta * uo + w2 * (same_base ? uo : bv[n]);
          slot = 1.0 * avg + bdt * dudt;  // the new state waits here until the next plane
          r0[n] = uo;  // the ring moves up one plane
          r1[n] = r2[n];
          r2[n] = r3[n];
          r3[n] = r4[n];
          // one component at a time: interleaving the seven reconstructions needs more
          // registers than the budget holds
          __builtin_amdgcn_sched_barrier(0);
        }
      };
      if (k == N - 1 && oz1 >= 0)
        zloop(std::true_type());
      else
        zloop(std::false_type());
    }
    flush(b, k1 - 1);  // the segment's last plane (own slots only: no barrier needed)
  }
  if (!partial) return;
  m = WaveMin(m);
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = wmin[0];
#pragma unroll
    for (int w = 1; w < NT / 64; ++w) v = DMin(v, wmin[w]);
    partial[blockIdx.x] = v;
  }
}
}  // namespace

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
                        int *groups, hipStream_t s, const int *list, int nlist) {
  if (!BurgersSweepSupported(t, V) || waves < 2 || waves > 4 || spread < 1 || spread > 3) {
    SetError("block sweep: unsupported block shape or component count");
    return PTH_FAIL;
  }
  const int N = t.fine.n[0];
  if (list && !(N == 16 && t.fine.g[0] == 4 && nlist > 0)) {
    SetError("block sweep: block lists need 16^3 blocks with 4 ghosts");
    return PTH_FAIL;
  }
  ProfScope prof("block_sweep", s);
  const int nb = list ? nlist : static_cast<int>(t.blocks.size());
  if (N == 16 && t.fine.g[0] == 4) {
    // register-ring kernel, compiled for 2 or 3 waves per SIMD
    const int minw = waves >= 3 ? 3 : 2;
    const int *fover16 = fc ? fc->dev_face : nullptr;
    const double *fcor16 = fc ? fc->store : nullptr;
    int grid16 = 0;
#define PTH_S16_R(NV, W, RC)                                                                   \
  {                                                                                            \
    const int r =                                                                              \
        ResidentGroups(reinterpret_cast<const void *>(&BlockSweep16Kernel<NV, W, RC>), 256);    \
    if (r < 1) {                                                                               \
      SetError("block sweep: occupancy query failed");                                         \
      return PTH_FAIL;                                                                         \
    }                                                                                          \
    grid16 = static_cast<int>(std::min<std::int64_t>(std::min(std::max(r / spread, 1),          \
                                                              max_partials),                   \
                                                     static_cast<std::int64_t>(nb) * 16));     \
    hipLaunchKernelGGL((BlockSweep16Kernel<NV, W, RC>), dim3(grid16), dim3(256), 0, s, U, base, \
                       out, derived, partial, coef, nb, beta, dtvar, fover16, fcor16, push,   \
                       list);                                                                  \
  }
#define PTH_S16_W(NV, W)                                                                       \
  if (recon == 0)                                                                              \
    PTH_S16_R(NV, W, 0)                                                                        \
  else                                                                                         \
    PTH_S16_R(NV, W, 1)
#define PTH_S16_CASE(NV)                                                                       \
  case NV:                                                                                     \
    if (minw == 3)                                                                             \
      PTH_S16_W(NV, 3)                                                                         \
    else                                                                                       \
      PTH_S16_W(NV, 2)                                                                         \
    break;
    switch (V) {
      PTH_S16_CASE(4)
      PTH_S16_CASE(5)
      PTH_S16_CASE(6)
      PTH_S16_CASE(7)
      PTH_S16_CASE(8)
      PTH_S16_CASE(9)
      PTH_S16_CASE(10)
      PTH_S16_CASE(11)
      PTH_S16_CASE(12)
      PTH_S16_CASE(13)
    }
#undef PTH_S16_CASE
#undef PTH_S16_W
#undef PTH_S16_R
    PTH_HIP(hipGetLastError());
    if (groups) *groups = grid16;
    return PTH_COMPLETE;
  }
  if (waves == 2) waves = 3;
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
  const bool ring = t.fine.n[0] == 16 && t.fine.g[0] == 4;
#define PTH_FD_CASE(NV)                                                                       \
  case NV:                                                                                    \
    if (ring)                                                                                 \
      hipLaunchKernelGGL((FlxcorDirectKernel<NV, true>), grid, dim3(64), 0, s, fc.dev, U,     \
                         fc.store, G, recon);                                                 \
    else                                                                                      \
      hipLaunchKernelGGL((FlxcorDirectKernel<NV, false>), grid, dim3(64), 0, s, fc.dev, U,    \
                         fc.store, G, recon);                                                 \
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

}  // namespace pth
