// Stencil helpers shared by the Burgers kernel files: cell and face operators of the
// reference's CalculateFluxes and flux correction, plus brick addressing.
//
// Everything here sits in an unnamed namespace, so each kernel file compiles its own copy
// with its own floating-point flags (pth_burgers.hip forbids FMA contraction, pth_sweep.hip
// allows it) and no definition is shared between the two.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pth_internal.hpp"

namespace pth {
namespace {

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

__device__ __forceinline__ double WaveMin(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = DMin(v, __shfl_xor(v, off, 64));
  return v;
}

// Loads and stores at 32-bit byte offsets from a per-workgroup base pointer.
__device__ __forceinline__ double LoadAt(const double *base, unsigned byte_off) {
  return *reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + byte_off);
}
__device__ __forceinline__ void StoreAt(double *base, unsigned byte_off, double v) {
  *reinterpret_cast<double *>(reinterpret_cast<char *>(base) + byte_off) = v;
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

// Brick shape of the topology's fine blocks.
inline Geo MakeGeo(const Topology &t, int V) {
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
}  // namespace pth
