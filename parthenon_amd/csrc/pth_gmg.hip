// Geometric-multigrid Poisson solver on the GPU (the reference's poisson_gmg MGSolver,
// src/solvers/mg_solver.hpp:37-521, with the PoissonEquation operator of
// example/poisson_gmg/poisson_equation.hpp:41-205), for uniform N^3 cell-centred grids on
// [0, 1]^3 with FixedFace-0 boundaries (boundary_conditions_generic.hpp:174-255).
//
// Mixed precision by iterative refinement:
//   outer (FP64):  r = b - A u  (with u <- u + e fused in),  ||r|| / ||b|| stop test
//   inner (T):     one V-cycle on A e = r from e = 0: weighted Jacobi pre-smoothing,
//                  residual + average restriction, coarser levels, linear prolongation of the
//                  correction, post-smoothing; the coarsest level only smooths
// T = float is the mixed solver, T = double the plain FP64 solver: same kernels.
//
// Levels are single arrays of n^3 cells (the blocks of the reference's GMG grids are a
// distribution artefact: on one GPU every level fits whole), so no level needs a ghost
// exchange; the Dirichlet ghost (2 * 0 - mirror) is formed on the fly at the domain faces.
// Levels of at most kBottom cells per side run inside one workgroup in LDS.
//
// Every kernel is batched: its grid covers a table of entries, one per (problem, level), so
// the same V-cycle stage of several concurrent solves is one launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "pth_device.hpp"
#include "pth_gmg.hpp"

namespace pth {
namespace {

constexpr int kWG = 256;

template <typename T>
__device__ __forceinline__ T Ratio(T a, T b) {  // robust::ratio (utils/robust.hpp:59-62)
  const T sgn = b >= T(0) ? T(1) : T(-1);
  const T small = T(10) * (sizeof(T) == 8 ? T(DBL_MIN) : T(FLT_MIN));
  return a / (b + sgn * small);
}

template <typename T>
__device__ __forceinline__ T Diagonal(T dx) {  // SetDiagonal (poisson_equation.hpp:86-125)
  const T D = T(1), alpha = T(0);
  T d = -(D + D) / (dx * dx) - alpha;
  d -= (D + D) / (dx * dx);
  d -= (D + D) / (dx * dx);
  return d;
}

// Per-level constants. Divisions by dx, by the restriction's total volume and by the coarse
// spacing become multiplications by their inverses where that gives the same bits (n a power
// of two: the divisors are powers of two), and always in FP32, where the inner cycle only has
// to reduce the error; the FP64 Jacobi keeps the reference's division by the diagonal.
template <typename T>
struct Lv {
  T dx, idx, D, rD, tv, itv;
  bool mul;
};
template <typename T>
__device__ __forceinline__ Lv<T> MakeLv(int n) {
  Lv<T> o;
  o.dx = T(1) / T(n);
  o.idx = T(n);
  o.mul = sizeof(T) == 4 || (n & (n - 1)) == 0;
  o.D = Diagonal(o.dx);
  const T small = T(10) * (sizeof(T) == 8 ? T(DBL_MIN) : T(FLT_MIN));
  o.rD = T(1) / (o.D + (o.D >= T(0) ? T(1) : T(-1)) * small);
  const T vol = o.dx * o.dx * o.dx;
  o.tv = ((vol + vol) + (vol + vol)) + ((vol + vol) + (vol + vol));
  o.itv = T(1) / o.tv;
  return o;
}

// 7-point operator at one cell in the reference's flux form (D = 1, alpha = 0); neighbours
// across a domain face are the FixedFace-0 ghost -u.
template <typename T>
__device__ __forceinline__ T ApplyA(T u, T xm, T xp, T ym, T yp, T zm, T zp, const Lv<T> &o) {
  const T D = T(1), alpha = T(0);
  T out = -alpha * u;
  if (o.mul) {  // D / dx == idx and x / dx == x * idx
    const T f1l = o.idx * (xm - u);
    const T f1r = o.idx * (u - xp);
    out += (f1l - f1r) * o.idx;
    const T f2l = D * (ym - u) * o.idx;
    const T f2r = D * (u - yp) * o.idx;
    out += (f2l - f2r) * o.idx;
    const T f3l = D * (zm - u) * o.idx;
    const T f3r = D * (u - zp) * o.idx;
    out += (f3l - f3r) * o.idx;
    return out;
  }
  const T dx = o.dx;
  const T f1l = D / dx * (xm - u);
  const T f1r = D / dx * (u - xp);
  out += (f1l - f1r) / dx;
  const T f2l = D * (ym - u) / dx;
  const T f2r = D * (u - yp) / dx;
  out += (f2l - f2r) / dx;
  const T f3l = D * (zm - u) / dx;
  const T f3r = D * (u - zp) / dx;
  out += (f3l - f3r) / dx;
  return out;
}

struct Cell {
  int i, j, k;
};
__device__ __forceinline__ Cell Decode(int c, int n) {
  Cell q;
  if ((n & (n - 1)) == 0) {  // power of two: shifts and masks (n is uniform per workgroup)
    const int sh = __builtin_ctz(static_cast<unsigned>(n));
    q.i = c & (n - 1);
    q.j = (c >> sh) & (n - 1);
    q.k = c >> (2 * sh);
    return q;
  }
  q.i = c % n;
  const int r = c / n;
  q.j = r % n;
  q.k = r / n;
  return q;
}

// the 6 neighbours of cell c of an n^3 array with the Dirichlet ghost
template <typename T>
struct Nbrs {
  T u, xm, xp, ym, yp, zm, zp;
};
template <typename T>
__device__ __forceinline__ Nbrs<T> Load(const T *__restrict__ x, int c, const Cell &q, int n) {
  Nbrs<T> v;
  const int s1 = n, s2 = n * n;
  v.u = x[c];
  v.xm = q.i > 0 ? x[c - 1] : -v.u;
  v.xp = q.i < n - 1 ? x[c + 1] : -v.u;
  v.ym = q.j > 0 ? x[c - s1] : -v.u;
  v.yp = q.j < n - 1 ? x[c + s1] : -v.u;
  v.zm = q.k > 0 ? x[c - s2] : -v.u;
  v.zp = q.k < n - 1 ? x[c + s2] : -v.u;
  return v;
}

// The same from an array in global memory, every load unconditional: a neighbour outside the
// array reads the cell itself and the ghost value -u replaces it afterwards. (gfx9 counts
// loads and stores on one in-order counter; loads under a per-lane branch make the compiler
// wait for each before the next.)
template <typename T>
__device__ __forceinline__ Nbrs<T> LoadG(const T *x0, int c, const Cell &q, int n) {
  const auto *x = Glob(x0);
  const int s1 = n, s2 = n * n;
  const T u = x[c];
  const T xm = x[q.i > 0 ? c - 1 : c], xp = x[q.i < n - 1 ? c + 1 : c];
  const T ym = x[q.j > 0 ? c - s1 : c], yp = x[q.j < n - 1 ? c + s1 : c];
  const T zm = x[q.k > 0 ? c - s2 : c], zp = x[q.k < n - 1 ? c + s2 : c];
  Nbrs<T> v;
  v.u = u;
  v.xm = q.i > 0 ? xm : -u;
  v.xp = q.i < n - 1 ? xp : -u;
  v.ym = q.j > 0 ? ym : -u;
  v.yp = q.j < n - 1 ? yp : -u;
  v.zm = q.k > 0 ? zm : -u;
  v.zp = q.k < n - 1 ? zp : -u;
  return v;
}

template <typename T>
__device__ __forceinline__ T JacobiPoint(const Nbrs<T> &v, T b, const Lv<T> &o, T w) {
  const T off = ApplyA(v.u, v.xm, v.xp, v.ym, v.yp, v.zm, v.zp, o) - o.D * v.u;
  const T val = b - off;
  const T q = sizeof(T) == 4 ? val * o.rD : Ratio(val, o.D);
  return w * q + (T(1) - w) * v.u;
}

// Entry q of a launch. Inline entries are selected with constant indices only, so the kernel
// arguments are read in place (a runtime index into them would copy them to scratch).
template <typename T>
__device__ __forceinline__ GmgEntry<T> At(const GmgArgs<T> &A, int q) {
  if (A.dev) return A.dev[q];
  GmgEntry<T> e = A.e[0];
#pragma unroll
  for (int i = 1; i < kGmgInline; ++i)
    if (i == q) e = A.e[i];
  return e;
}

// The entry whose workgroup range holds this workgroup: entries sorted by wg_begin, the last
// one ends at gridDim.x. The logical workgroup id is XCD-remapped, so each XCD walks
// contiguous rows and planes of a level and finds the y and z neighbours of its rows in its
// own L2.
template <typename T>
__device__ __forceinline__ GmgEntry<T> Find(const GmgArgs<T> &A, int &local) {
  const int w = static_cast<int>(XcdRemap(blockIdx.x, gridDim.x));
  GmgEntry<T> e;
  if (A.dev) {  // binary search over the staged entries
    int lo = 0, hi = A.ne - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (A.dev[mid].wg_begin <= w)
        lo = mid;
      else
        hi = mid - 1;
    }
    e = A.dev[lo];
  } else {
    e = A.e[0];
#pragma unroll
    for (int i = 1; i < kGmgInline; ++i)
      if (i < A.ne && A.e[i].wg_begin <= w) e = A.e[i];
  }
  local = w - e.wg_begin;
  return e;
}

// ---- kernels --------------------------------------------------------------------------

// FP64 outer step: u_new = u_old + e (e may be null), r = b - A u_new, r stored in T as the
// inner right-hand side, per-workgroup partial sums of r^2.
// One cell of the outer step; CORR: a correction is pending (e.x). Every load comes before
// the first store, so a lane has its 8 (or 15) loads in flight at once.
template <typename T, bool CORR>
__device__ __forceinline__ double OuterCell(const GmgEntry<T> &e, int c, const Cell &q, int n) {
  const Lv<double> o = MakeLv<double>(n);
  const double bv = Glob(static_cast<const double *>(e.b64))[c];
  Nbrs<double> v = LoadG(static_cast<const double *>(e.u_old), c, q, n);
  if (CORR) {  // add the correction at the 7 points (the ghost follows the updated value)
    const Nbrs<T> d = LoadG(static_cast<const T *>(e.x), c, q, n);
    v.u += static_cast<double>(d.u);
    v.xm = q.i > 0 ? v.xm + static_cast<double>(d.xm) : -v.u;
    v.xp = q.i < n - 1 ? v.xp + static_cast<double>(d.xp) : -v.u;
    v.ym = q.j > 0 ? v.ym + static_cast<double>(d.ym) : -v.u;
    v.yp = q.j < n - 1 ? v.yp + static_cast<double>(d.yp) : -v.u;
    v.zm = q.k > 0 ? v.zm + static_cast<double>(d.zm) : -v.u;
    v.zp = q.k < n - 1 ? v.zp + static_cast<double>(d.zp) : -v.u;
    Glob(e.u_new)[c] = v.u;
  }
  const double r = bv - ApplyA(v.u, v.xm, v.xp, v.ym, v.yp, v.zm, v.zp, o);
  Glob(e.b)[c] = static_cast<T>(r);
  return r;
}

template <typename T>
__global__ void __launch_bounds__(kWG) OuterKernel(const GmgArgs<T> A) {
  __shared__ double wsum[kWG / 64];
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, c = local * kWG + threadIdx.x;
  double rr = 0.0;
  if (c < n * n * n) {
    const Cell q = Decode(c, n);
    const double r = e.x ? OuterCell<T, true>(e, c, q, n) : OuterCell<T, false>(e, c, q, n);
    rr = r * r;
  }
  for (int off = 32; off > 0; off >>= 1) rr += __shfl_xor(rr, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = rr;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kWG / 64; ++w) s += wsum[w];
    e.partial[local] = s;
  }
}

// per-entry sum of the partials of OuterKernel (or of the rhs norm): one workgroup of kNormWG
// threads per entry, each thread adding four interleaved streams of partials into separate
// sums (independent loads in flight), then the workgroup tree; per_wg = cells per workgroup
// of the kernel that wrote the partials. The order is fixed, so an entry's norm is the same
// alone and in a batch.
constexpr int kNormWG = 1024;
template <typename T>
__global__ void __launch_bounds__(kNormWG)
    NormKernel(const GmgArgs<T> A, double *__restrict__ out, int per_wg) {
  __shared__ double wsum[kNormWG / 64];
  const GmgEntry<T> e = At(A, blockIdx.x);
  const int n = e.n, np = (n * n * n + per_wg - 1) / per_wg;
  const double *pa = e.partial;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int p = threadIdx.x; p < np; p += 4 * kNormWG) {
    s0 += pa[p];
    if (p + kNormWG < np) s1 += pa[p + kNormWG];
    if (p + 2 * kNormWG < np) s2 += pa[p + 2 * kNormWG];
    if (p + 3 * kNormWG < np) s3 += pa[p + 3 * kNormWG];
  }
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kNormWG / 64; ++w) t += wsum[w];
    out[blockIdx.x] = t;
  }
}

// partial sums of b^2 (the stop test's denominator)
template <typename T>
__global__ void __launch_bounds__(kWG) RhsNormKernel(const GmgArgs<T> A) {
  __shared__ double wsum[kWG / 64];
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, c = local * kWG + threadIdx.x;
  double s = 0.0;
  if (c < n * n * n) s = Glob(e.b64)[c] * Glob(e.b64)[c];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kWG / 64; ++w) t += wsum[w];
    e.partial[local] = t;
  }
}

// Pre-smoothing from x = 0: one sweep (PRE2 = false) or two sweeps fused, the first sweep's
// values at the 7 points being formed from b where they are needed (the same operations as
// two separate sweeps, so the same bits).
template <typename T, bool PRE2>
__global__ void __launch_bounds__(kWG)
    PreZeroKernel(const GmgArgs<T> A, T w) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, c = local * kWG + threadIdx.x;
  if (c >= n * n * n) return;
  const Cell q = Decode(c, n);
  const Lv<T> o = MakeLv<T>(n);
  const Nbrs<T> zero = {T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  const T *b = e.b;
  if (!PRE2) {
    Glob(e.x)[c] = JacobiPoint(zero, Glob(b)[c], o, w);
    return;
  }
  const Nbrs<T> bb = LoadG(b, c, q, n);
  Nbrs<T> x1;
  x1.u = JacobiPoint(zero, bb.u, o, w);
  x1.xm = q.i > 0 ? JacobiPoint(zero, bb.xm, o, w) : -x1.u;
  x1.xp = q.i < n - 1 ? JacobiPoint(zero, bb.xp, o, w) : -x1.u;
  x1.ym = q.j > 0 ? JacobiPoint(zero, bb.ym, o, w) : -x1.u;
  x1.yp = q.j < n - 1 ? JacobiPoint(zero, bb.yp, o, w) : -x1.u;
  x1.zm = q.k > 0 ? JacobiPoint(zero, bb.zm, o, w) : -x1.u;
  x1.zp = q.k < n - 1 ? JacobiPoint(zero, bb.zp, o, w) : -x1.u;
  Glob(e.x)[c] = JacobiPoint(x1, bb.u, o, w);
}

template <typename T>
__global__ void __launch_bounds__(kWG) ZeroKernel(const GmgArgs<T> A) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int c = local * kWG + threadIdx.x;
  if (c < e.n * e.n * e.n) Glob(e.x)[c] = T(0);
}

// one Jacobi sweep x -> t (the caller swaps x and t)
template <typename T>
__global__ void __launch_bounds__(kWG)
    JacobiKernel(const GmgArgs<T> A, T w) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, c = local * kWG + threadIdx.x;
  if (c >= n * n * n) return;
  const Cell q = Decode(c, n);
  const Lv<T> o = MakeLv<T>(n);
  Glob(e.t)[c] = JacobiPoint(LoadG(static_cast<const T *>(e.x), c, q, n), Glob(e.b)[c], o, w);
}

// One Jacobi sweep x -> t by z columns: a workgroup covers 64 x 4 cells of NZ consecutive
// planes, each thread marches its column keeping the z neighbours in registers (the x and y
// neighbours come from the cache). Same operations as JacobiKernel; for n % 64 == 0 and
// n % NZ == 0.
template <typename T, int NZ>
__global__ void __launch_bounds__(kWG) JacobiZKernel(const GmgArgs<T> A, T w) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, tx = n / 64, ty = n / 4;
  const int bx = local % tx, r = local / tx, by = r % ty, bz = r / ty;
  const int i = bx * 64 + static_cast<int>(threadIdx.x % 64);
  const int j = by * 4 + static_cast<int>(threadIdx.x / 64);
  const int k0 = bz * NZ;
  const Lv<T> o = MakeLv<T>(n);
  const int s1 = n, s2 = n * n;
  const auto *x = Glob(static_cast<const T *>(e.x));
  const auto *b = Glob(static_cast<const T *>(e.b));
  auto *t = Glob(e.t);
  int c = (k0 * n + j) * n + i;
  // neighbours outside the array read the cell itself (unconditional loads) and are replaced
  T cur = x[c];
  T prev = x[k0 > 0 ? c - s2 : c];
  T next = x[k0 + 1 < n ? c + s2 : c];
#pragma unroll 2
  for (int k = k0; k < k0 + NZ; ++k, c += s2) {
    const T xm = x[i > 0 ? c - 1 : c], xp = x[i < n - 1 ? c + 1 : c];
    const T ym = x[j > 0 ? c - s1 : c], yp = x[j < n - 1 ? c + s1 : c];
    const T bc = b[c];
    const T nn = x[k + 2 < n ? c + 2 * s2 : c];
    Nbrs<T> v;
    v.u = cur;
    v.xm = i > 0 ? xm : -v.u;
    v.xp = i < n - 1 ? xp : -v.u;
    v.ym = j > 0 ? ym : -v.u;
    v.yp = j < n - 1 ? yp : -v.u;
    v.zm = k > 0 ? prev : -v.u;
    v.zp = k < n - 1 ? next : -v.u;
    t[c] = JacobiPoint(v, bc, o, w);
    prev = cur;
    cur = next;
    next = nn;
  }
}

// Column form of the fused pre-smoothing: a workgroup covers 64 x 4 cells of NZ consecutive
// planes and each thread marches its column with the z neighbours of b in registers; the
// operations per cell are those of PreZeroKernel<T, true>. (A column form of the FP64 outer
// step measured slower than OuterKernel and is not used.)
struct ZCell {
  int i, j, k0, c;
};
__device__ __forceinline__ ZCell ZStart(int n, int local, int nz) {
  const int tx = n / 64, ty = n / 4;
  const int bx = local % tx, r = local / tx, by = r % ty, bz = r / ty;
  ZCell z;
  z.i = bx * 64 + static_cast<int>(threadIdx.x % 64);
  z.j = by * 4 + static_cast<int>(threadIdx.x / 64);
  z.k0 = bz * nz;
  z.c = (z.k0 * n + z.j) * n + z.i;
  return z;
}

template <typename T, int NZ>
__global__ void __launch_bounds__(kWG) PreZero2ZKernel(const GmgArgs<T> A, T w) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, s1 = n, s2 = n * n;
  const ZCell z = ZStart(n, local, NZ);
  const int i = z.i, j = z.j;
  int c = z.c;
  const Lv<T> o = MakeLv<T>(n);
  const Nbrs<T> zero = {T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  const auto *b = Glob(static_cast<const T *>(e.b));
  auto *x = Glob(e.x);
  // neighbours outside the array read the cell itself (unconditional loads) and are replaced
  T bprev = b[z.k0 > 0 ? c - s2 : c], bcur = b[c];
  T bnext = b[z.k0 + 1 < n ? c + s2 : c];
  for (int k = z.k0; k < z.k0 + NZ; ++k, c += s2) {
    const T bxm = b[i > 0 ? c - 1 : c], bxp = b[i < n - 1 ? c + 1 : c];
    const T bym = b[j > 0 ? c - s1 : c], byp = b[j < n - 1 ? c + s1 : c];
    const T bnn = b[k + 2 < n ? c + 2 * s2 : c];
    Nbrs<T> x1;
    x1.u = JacobiPoint(zero, bcur, o, w);
    x1.xm = i > 0 ? JacobiPoint(zero, bxm, o, w) : -x1.u;
    x1.xp = i < n - 1 ? JacobiPoint(zero, bxp, o, w) : -x1.u;
    x1.ym = j > 0 ? JacobiPoint(zero, bym, o, w) : -x1.u;
    x1.yp = j < n - 1 ? JacobiPoint(zero, byp, o, w) : -x1.u;
    x1.zm = k > 0 ? JacobiPoint(zero, bprev, o, w) : -x1.u;
    x1.zp = k < n - 1 ? JacobiPoint(zero, bnext, o, w) : -x1.u;
    x[c] = JacobiPoint(x1, bcur, o, w);
    bprev = bcur;
    bcur = bnext;
    bnext = bnn;
  }
}

template <typename T>
__device__ __forceinline__ T Restrict8(const T t[2][2][2], const Lv<T> &o) {
  const T sum = ((t[0][0][0] + t[0][1][0]) + (t[0][0][1] + t[0][1][1])) +
                ((t[1][0][0] + t[1][1][0]) + (t[1][0][1] + t[1][1][1]));
  return o.mul ? sum * o.itv : sum / o.tv;
}

// residual of the 8 children of one coarse cell and their average (RestrictAverage) into the
// coarser level's rhs; one thread per coarse cell
template <typename T>
__global__ void __launch_bounds__(kWG)
    ResidualRestrictKernel(const GmgArgs<T> A) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int nc = e.nc, n = e.n, cc = local * kWG + threadIdx.x;
  if (cc >= nc * nc * nc) return;
  const Cell p = Decode(cc, nc);
  const Lv<T> o = MakeLv<T>(n);
  const T vol = o.dx * o.dx * o.dx;
  T t[2][2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int bb = 0; bb < 2; ++bb)
#pragma unroll
      for (int c2 = 0; c2 < 2; ++c2) {
        const Cell q = {2 * p.i + c2, 2 * p.j + bb, 2 * p.k + a};
        const int c = (q.k * n + q.j) * n + q.i;
        const Nbrs<T> v = LoadG(static_cast<const T *>(e.x), c, q, n);
        const T r = Glob(e.b)[c] - ApplyA(v.u, v.xm, v.xp, v.ym, v.yp, v.zm, v.zp, o);
        t[a][bb][c2] = vol * r;
      }
  Glob(e.bc)[cc] = Restrict8(t, o);
}

// ProlongateSharedLinear of the coarse correction xc at fine cell q of an n^3 level (coarse
// ghosts from the Dirichlet condition; pr_ops.hpp:166-289 term grouping). Every load is
// unconditional: a coarse neighbour outside the level reads the cell itself and the ghost
// value replaces it.
template <typename T, typename P>
__device__ __forceinline__ T ProlongValue(const P *xc, int nc, const Cell &q) {
  const Cell p = {q.i >> 1, q.j >> 1, q.k >> 1};
  const int oi = q.i & 1, oj = q.j & 1, ok = q.k & 1;
  const int cc = (p.k * nc + p.j) * nc + p.i;
  const T fc = xc[cc];
  const Lv<T> oc = MakeLv<T>(nc);
  const T hc = oc.dx, df = T(0.25) * hc;
  auto grad = [&](T diff) { return oc.mul ? diff * oc.idx : diff / hc; };
  // the one-sided gradient on this child's side of the coarse cell, per dimension
  const bool in1 = oi ? p.i < nc - 1 : p.i > 0;
  const bool in2 = oj ? p.j < nc - 1 : p.j > 0;
  const bool in3 = ok ? p.k < nc - 1 : p.k > 0;
  const T l1 = xc[in1 ? cc + (oi ? 1 : -1) : cc];
  const T l2 = xc[in2 ? cc + (oj ? nc : -nc) : cc];
  const T l3 = xc[in3 ? cc + (ok ? nc * nc : -nc * nc) : cc];
  const T f1 = in1 ? l1 : -fc;
  const T f2 = in2 ? l2 : -fc;
  const T f3 = in3 ? l3 : -fc;
  const T g1 = grad(oi ? f1 - fc : fc - f1);
  const T g2 = grad(oj ? f2 - fc : fc - f2);
  const T g3 = grad(ok ? f3 - fc : fc - f3);
  // fc +/- (g1 df +/- g2 df +/- g3 df) with the signs of the child's offsets, grouped like the
  // reference's eight cases (pr_ops.hpp:261-283): the outer sign is the x offset's, the inner
  // signs are relative to it; a product with -1 and an added negation round exactly like the
  // subtractions they stand for
  const T so = oi ? T(1) : T(-1);
  const T sb = oi == oj ? T(1) : T(-1);
  const T sc = oi == ok ? T(1) : T(-1);
  return fc + so * ((g1 * df + sb * (g2 * df)) + sc * (g3 * df));
}

// x += ProlongateSharedLinear(coarse x); one thread per fine cell
template <typename T>
__global__ void __launch_bounds__(kWG) ProlongKernel(const GmgArgs<T> A) {
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, c = local * kWG + threadIdx.x;
  if (c >= n * n * n) return;
  const Cell q = Decode(c, n);
  const T v = ProlongValue<T>(Glob(static_cast<const T *>(e.xc)), e.nc, q);
  Glob(e.x)[c] += v;
}

// Prolongation fused into the first post-smoothing sweep, column form: t = J(x + P xc), the
// same values as ProlongKernel followed by JacobiZKernel. A workgroup covers 64 x 4 cells of
// NZ consecutive planes; per plane, the corrected values x' = x + P xc of its cells and of the
// one-cell ring around them (rows j0 - 1 and j0 + 4, columns i0 - 1 and i0 + 64) are staged
// in LDS for the x and y neighbours, and each thread keeps x' of its column at k - 1, k, k + 1
// in registers. The level's x is read, never written (after the launch t holds the iterate).
template <typename T, int NZ>
__global__ void __launch_bounds__(kWG) ProlongJacobiZKernel(const GmgArgs<T> A, T w) {
  constexpr int TW = 66;  // tile width: 64 cells and one on each side
  __shared__ T tile[6][TW];
  int local;
  const GmgEntry<T> e = Find(A, local);
  const int n = e.n, nc = e.nc, s2 = n * n;
  const ZCell z = ZStart(n, local, NZ);
  const int i = z.i, j = z.j, lx = static_cast<int>(threadIdx.x % 64),
            ly = static_cast<int>(threadIdx.x / 64);
  const int i0 = i - lx, j0 = j - ly;
  int c = z.c;
  const Lv<T> o = MakeLv<T>(n);
  const auto *x = Glob(static_cast<const T *>(e.x));
  const auto *b = Glob(static_cast<const T *>(e.b));
  const auto *xc = Glob(static_cast<const T *>(e.xc));
  auto *t = Glob(e.t);
  // x + P xc at a cell of the level (callers clamp positions into the level)
  auto xp1 = [&](int ii, int jj, int kk) -> T {
    const Cell q = {ii, jj, kk};
    return x[(kk * n + jj) * n + ii] + ProlongValue<T>(xc, nc, q);
  };
  // the ring cell this thread fills: 66 + 66 cells of rows 0 and 5, then columns 0 and 65 of
  // rows 1..4 (140 cells; threads 140..255 fill none)
  const int h = static_cast<int>(threadIdx.x);
  const int hr = h < TW ? 0 : h < 2 * TW ? 5 : 1 + (h - 2 * TW) / 2;
  const int hc = h < TW ? h : h < 2 * TW ? h - TW : ((h - 2 * TW) % 2 ? TW - 1 : 0);
  const int hi = min(max(i0 - 1 + hc, 0), n - 1), hj = min(max(j0 - 1 + hr, 0), n - 1);
  const bool ring = h < 2 * TW + 8;
  T prev = xp1(i, j, max(z.k0 - 1, 0));
  T cur = xp1(i, j, z.k0);
  T next = xp1(i, j, min(z.k0 + 1, n - 1));
  for (int k = z.k0; k < z.k0 + NZ; ++k, c += s2) {
    tile[ly + 1][lx + 1] = cur;
    if (ring) tile[hr][hc] = xp1(hi, hj, k);
    const T bc = b[c];
    const T nn = xp1(i, j, min(k + 2, n - 1));
    __syncthreads();
    Nbrs<T> v;
    v.u = cur;
    v.xm = i > 0 ? tile[ly + 1][lx] : -v.u;
    v.xp = i < n - 1 ? tile[ly + 1][lx + 2] : -v.u;
    v.ym = j > 0 ? tile[ly][lx + 1] : -v.u;
    v.yp = j < n - 1 ? tile[ly + 2][lx + 1] : -v.u;
    v.zm = k > 0 ? prev : -v.u;
    v.zp = k < n - 1 ? next : -v.u;
    t[c] = JacobiPoint(v, bc, o, w);
    __syncthreads();  // the next plane overwrites the tile
    prev = cur;
    cur = next;
    next = nn;
  }
}

// ---- bottom of the V-cycle in one workgroup (levels of at most NB cells per side) --------
template <typename T>
__device__ void BSweep(T *x, T *t, const T *b, int n, T w, bool zero) {
  const int N3 = n * n * n;
  const Lv<T> o = MakeLv<T>(n);
  const Nbrs<T> zn = {T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int c = threadIdx.x; c < N3; c += blockDim.x) {
    const Cell q = Decode(c, n);
    t[c] = JacobiPoint(zero ? zn : Load(static_cast<const T *>(x), c, q, n), b[c], o, w);
  }
  __syncthreads();
  for (int c = threadIdx.x; c < N3; c += blockDim.x) x[c] = t[c];
  __syncthreads();
}

template <typename T>
__device__ void BRestrict(const T *x, const T *b, T *bc, T *xc, int n) {
  const int nc = n / 2, N3 = nc * nc * nc;
  const Lv<T> o = MakeLv<T>(n);
  const T vol = o.dx * o.dx * o.dx;
  for (int cc = threadIdx.x; cc < N3; cc += blockDim.x) {
    const Cell p = Decode(cc, nc);
    T t[2][2][2];
    for (int a = 0; a < 2; ++a)
      for (int bb = 0; bb < 2; ++bb)
        for (int c2 = 0; c2 < 2; ++c2) {
          const Cell q = {2 * p.i + c2, 2 * p.j + bb, 2 * p.k + a};
          const int c = (q.k * n + q.j) * n + q.i;
          const Nbrs<T> v = Load(x, c, q, n);
          t[a][bb][c2] = vol * (b[c] - ApplyA(v.u, v.xm, v.xp, v.ym, v.yp, v.zm, v.zp, o));
        }
    bc[cc] = Restrict8(t, o);
    xc[cc] = T(0);
  }
  __syncthreads();
}

template <typename T>
__device__ void BProlong(T *x, const T *xc, int n) {
  const int nc = n / 2, N3 = n * n * n;
  const Lv<T> oc = MakeLv<T>(nc);
  const T hc = oc.dx, df = T(0.25) * hc;
  auto grad = [&](T diff) { return oc.mul ? diff * oc.idx : diff / hc; };
  for (int c = threadIdx.x; c < N3; c += blockDim.x) {
    const Cell q = Decode(c, n);
    const Cell p = {q.i >> 1, q.j >> 1, q.k >> 1};
    const int oi = q.i & 1, oj = q.j & 1, ok = q.k & 1;
    const int cc = (p.k * nc + p.j) * nc + p.i;
    const T fc = xc[cc];
    const T g1 = oi == 0 ? grad(fc - (p.i > 0 ? xc[cc - 1] : -fc))
                         : grad((p.i < nc - 1 ? xc[cc + 1] : -fc) - fc);
    const T g2 = oj == 0 ? grad(fc - (p.j > 0 ? xc[cc - nc] : -fc))
                         : grad((p.j < nc - 1 ? xc[cc + nc] : -fc) - fc);
    const T g3 = ok == 0 ? grad(fc - (p.k > 0 ? xc[cc - nc * nc] : -fc))
                         : grad((p.k < nc - 1 ? xc[cc + nc * nc] : -fc) - fc);
    T v;
    switch ((ok << 2) | (oj << 1) | oi) {
      case 0: v = fc - (g1 * df + g2 * df + g3 * df); break;
      case 1: v = fc + (g1 * df - g2 * df - g3 * df); break;
      case 2: v = fc - (g1 * df - g2 * df + g3 * df); break;
      case 3: v = fc + (g1 * df + g2 * df - g3 * df); break;
      case 4: v = fc - (g1 * df + g2 * df - g3 * df); break;
      case 5: v = fc + (g1 * df - g2 * df + g3 * df); break;
      case 6: v = fc - (g1 * df - g2 * df - g3 * df); break;
      default: v = fc + (g1 * df + g2 * df + g3 * df); break;
    }
    x[c] += v;
  }
  __syncthreads();
}

// One workgroup per entry: the V-cycle from this entry's level (rhs e.b, result in e.x) down
// to the coarsest level, every level in LDS.
template <typename T, int NB>
__global__ void __launch_bounds__(1024)
    BottomKernel(const GmgArgs<T> A, T w, int pre, int post) {
  constexpr int kCap = NB * NB * NB + (NB / 2) * (NB / 2) * (NB / 2) * 8 / 7 + 8;
  __shared__ T sx[kCap], st[kCap], sb[kCap];
  const GmgEntry<T> e = At(A, blockIdx.x);
  int nl = 0, off[8], nn[8];
  for (int n = e.n, o = 0; nl < 8; n /= 2) {
    off[nl] = o;
    nn[nl] = n;
    ++nl;
    o += n * n * n;
    if (n % 2 != 0 || n == 1) break;
  }
  const int N3 = e.n * e.n * e.n;
  const bool warm = e.warm != 0;
  for (int c = threadIdx.x; c < N3; c += blockDim.x) {
    sb[c] = e.b[c];
    if (warm) sx[c] = e.x[c];
  }
  __syncthreads();
  for (int l = 0; l < nl; ++l) {
    T *x = sx + off[l], *t = st + off[l], *b = sb + off[l];
    const bool from_x = warm && l == 0;
    for (int s = 0; s < pre; ++s) BSweep(x, t, b, nn[l], w, s == 0 && !from_x);
    if (pre == 0 && !from_x) {
      for (int c = threadIdx.x; c < nn[l] * nn[l] * nn[l]; c += blockDim.x) x[c] = T(0);
      __syncthreads();
    }
    if (l + 1 == nl) break;
    BRestrict(static_cast<const T *>(x), static_cast<const T *>(b), sb + off[l + 1],
              sx + off[l + 1], nn[l]);
  }
  for (int l = nl - 2; l >= 0; --l) {
    T *x = sx + off[l], *t = st + off[l], *b = sb + off[l];
    BProlong(x, static_cast<const T *>(sx + off[l + 1]), nn[l]);
    for (int s = 0; s < post; ++s) BSweep(x, t, b, nn[l], w, false);
  }
  for (int c = threadIdx.x; c < N3; c += blockDim.x) e.x[c] = sx[c];
}

// Right-hand side: 16 seeded Gaussian blobs at cell centres (FP64)
struct BlobParams {
  double v[80];  // per blob: cx, cy, cz, width, sign
};
__global__ void __launch_bounds__(kWG) BlobKernel(double *__restrict__ b, int n, BlobParams P) {
  const double *blob = P.v;
  const int c = blockIdx.x * kWG + threadIdx.x;
  if (c >= n * n * n) return;
  const Cell q = Decode(c, n);
  const double h = 1.0 / n;
  const double x = (q.i + 0.5) * h, y = (q.j + 0.5) * h, z = (q.k + 0.5) * h;
  double v = 0.0;
  for (int m = 0; m < 16; ++m) {
    const double *p = blob + 5 * m;
    const double dx = x - p[0], dy = y - p[1], dz = z - p[2];
    v += p[4] * exp(-(dx * dx + dy * dy + dz * dz) / (p[3] * p[3]));
  }
  b[c] = v;
}

__global__ void __launch_bounds__(kWG) FillKernel(double *__restrict__ a, std::int64_t n) {
  const std::int64_t c = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x;
  if (c < n) a[c] = 0.0;
}

int Blocks(std::int64_t cells) { return static_cast<int>((cells + kWG - 1) / kWG); }

constexpr int kZCol = 8;    // planes per thread of the column Jacobi sweep
constexpr int kPJCol = 8;   // planes per thread of the prolongation-and-sweep kernel

}  // namespace

// ---- host side -------------------------------------------------------------------------

void GmgBlobs(std::uint64_t seed, double out[80]) {
  std::uint64_t st = seed;
  auto next = [&]() {
    std::uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return static_cast<double>((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
  };
  for (int m = 0; m < 16; ++m) {
    for (int d = 0; d < 3; ++d) out[5 * m + d] = 0.1 + 0.8 * next();
    out[5 * m + 3] = 0.02 + 0.06 * next();
    out[5 * m + 4] = next() < 0.5 ? -1.0 : 1.0;
  }
}

std::vector<int> GmgLevelSizes(int N) {
  std::vector<int> s;
  for (int n = N;; n /= 2) {
    s.push_back(n);
    if (n % 2 != 0 || n == 1) break;
  }
  return s;
}

template <typename T>
std::int64_t GmgProblem<T>::Bytes(int N) {
  std::int64_t b = 0;
  for (int n : GmgLevelSizes(N)) b += 3 * static_cast<std::int64_t>(n) * n * n * sizeof(T);
  const std::int64_t c = static_cast<std::int64_t>(N) * N * N;
  return b + 3 * c * 8 + ((c + kWG - 1) / kWG) * 8;
}

template <typename T>
pth_status GmgProblem<T>::Allocate(int N_, int bottom) {
  N = N_;
  sizes = GmgLevelSizes(N);
  nbottom = 0;
  while (nbottom < static_cast<int>(sizes.size()) && sizes[nbottom] > bottom) ++nbottom;
  if (nbottom == static_cast<int>(sizes.size())) {
    SetError("GMG: the coarsest level is larger than the in-LDS bottom solver");
    return PTH_FAIL;
  }
  const std::int64_t c = static_cast<std::int64_t>(N) * N * N;
  if (c >= (std::int64_t(1) << 31)) {
    SetError("GMG: grid too large");
    return PTH_FAIL;
  }
  std::int64_t total = 0;
  for (int n : sizes) total += 3 * static_cast<std::int64_t>(n) * n * n;
  PTH_HIP(hipMalloc(&levels, sizeof(T) * total));
  PTH_HIP(hipMemset(levels, 0, sizeof(T) * total));
  PTH_HIP(hipMalloc(&u[0], sizeof(double) * c));
  PTH_HIP(hipMalloc(&u[1], sizeof(double) * c));
  PTH_HIP(hipMalloc(&b64, sizeof(double) * c));
  PTH_HIP(hipMalloc(&partial, sizeof(double) * ((c + kWG - 1) / kWG)));
  PTH_HIP(hipMemset(u[0], 0, sizeof(double) * c));
  PTH_HIP(hipMemset(u[1], 0, sizeof(double) * c));
  PTH_HIP(hipMemset(b64, 0, sizeof(double) * c));
  PTH_HIP(hipDeviceSynchronize());  // the clears above are ordered before any stream's work
  lv.clear();
  std::int64_t o = 0;
  for (int n : sizes) {
    const std::int64_t s = static_cast<std::int64_t>(n) * n * n;
    lv.push_back({levels + o, levels + o + s, levels + o + 2 * s});
    o += 3 * s;
  }
  cur = 0;
  return PTH_COMPLETE;
}

template <typename T>
void GmgProblem<T>::Free() {
  if (levels) (void)hipFree(levels);
  for (double *&p : u)
    if (p) (void)hipFree(p);
  if (b64) (void)hipFree(b64);
  if (partial) (void)hipFree(partial);
  levels = nullptr;
  u[0] = u[1] = b64 = partial = nullptr;
}

template <typename T>
GmgEntry<T> GmgProblem<T>::Entry(int l) const {
  GmgEntry<T> e;
  std::memset(&e, 0, sizeof(e));
  e.x = lv[l].x;
  e.t = lv[l].t;
  e.b = lv[l].b;
  if (l + 1 < static_cast<int>(lv.size())) {
    e.xc = lv[l + 1].x;
    e.bc = lv[l + 1].b;
    e.nc = sizes[l + 1];
  }
  e.n = sizes[l];
  e.u_old = u[cur];
  e.u_new = u[1 - cur];
  e.b64 = b64;
  e.partial = partial;
  return e;
}

template <typename T>
pth_status GmgBatch<T>::Stage(const std::vector<GmgEntry<T>> &es, bool coarse_cells,
                              GmgArgs<T> *args, int *nwg, int per_wg) {
  std::vector<GmgEntry<T>> v = es;
  int wg = 0;
  for (auto &e : v) {
    e.wg_begin = wg;
    const std::int64_t n = coarse_cells ? e.nc : e.n;
    wg += static_cast<int>((n * n * n + per_wg - 1) / per_wg);
  }
  std::memset(args, 0, sizeof(*args));
  args->ne = static_cast<int>(v.size());
  *nwg = wg;
  if (static_cast<int>(v.size()) <= kGmgInline) {  // in the kernel arguments
    std::copy(v.begin(), v.end(), args->e);
    return PTH_COMPLETE;
  }
  // larger batches: written to a host staging ring and copied once per stage
  if (static_cast<int>(v.size()) > kStageCap) {
    SetError("GMG: too many batched entries in one launch");
    return PTH_FAIL;
  }
  if (stage_used + static_cast<int>(v.size()) > kStageCap) {
    // the ring is full: the copies queued from it must have run before it is rewritten
    PTH_HIP(hipStreamSynchronize(stream));
    stage_used = 0;
  }
  GmgEntry<T> *h = stage_host + stage_used, *d = stage_dev + stage_used;
  std::copy(v.begin(), v.end(), h);
  PTH_HIP(hipMemcpyAsync(d, h, sizeof(GmgEntry<T>) * v.size(), hipMemcpyHostToDevice, stream));
  stage_used += static_cast<int>(v.size());
  args->dev = d;
  return PTH_COMPLETE;
}

template <typename T>
pth_status GmgBatch<T>::Init(hipStream_t s) {
  stream = s;
  if (!stage_host) {
    PTH_HIP(hipHostMalloc(&stage_host, sizeof(GmgEntry<T>) * kStageCap));
    PTH_HIP(hipMalloc(&stage_dev, sizeof(GmgEntry<T>) * kStageCap));
    PTH_HIP(hipHostMalloc(&norms_host, sizeof(double) * kMaxProblems));
    PTH_HIP(hipMalloc(&norms_dev, sizeof(double) * kMaxProblems));
  }
  return PTH_COMPLETE;
}

template <typename T>
void GmgBatch<T>::Free() {
  if (stage_host) (void)hipHostFree(stage_host);
  if (stage_dev) (void)hipFree(stage_dev);
  if (norms_host) (void)hipHostFree(norms_host);
  if (norms_dev) (void)hipFree(norms_dev);
  stage_host = nullptr;
  stage_dev = nullptr;
  norms_host = nullptr;
  norms_dev = nullptr;
}

template <typename T>
pth_status GmgBatch<T>::RhsNorms(const std::vector<GmgProblem<T> *> &ps, double *out) {
  std::vector<GmgEntry<T>> es;
  for (auto *p : ps) es.push_back(p->Entry(0));
  GmgArgs<T> a;
  int nwg;
  if (Stage(es, false, &a, &nwg) != PTH_COMPLETE) return PTH_FAIL;
  const int ne = a.ne;
  hipLaunchKernelGGL(RhsNormKernel<T>, dim3(nwg), dim3(kWG), 0, stream, a);
  hipLaunchKernelGGL(NormKernel<T>, dim3(ne), dim3(kNormWG), 0, stream, a, norms_dev, kWG);
  PTH_HIP(hipGetLastError());
  PTH_HIP(hipMemcpyAsync(norms_host, norms_dev, sizeof(double) * ne, hipMemcpyDeviceToHost,
                         stream));
  PTH_HIP(hipStreamSynchronize(stream));
  stage_used = 0;  // every queued copy has run
  for (int q = 0; q < ne; ++q) out[q] = std::sqrt(norms_host[q]);
  return PTH_COMPLETE;
}

// Outer FP64 step for every problem in ps: apply the pending correction (when correct[q]),
// form r = b - A u as the inner right-hand side, and return ||r|| in out[] (synchronizes).
template <typename T>
pth_status GmgBatch<T>::Residual(const std::vector<GmgProblem<T> *> &ps,
                                 const std::vector<char> &correct, double *out) {
  std::vector<GmgEntry<T>> es;
  for (size_t q = 0; q < ps.size(); ++q) {
    GmgEntry<T> e = ps[q]->Entry(0);
    if (!correct[q]) e.x = nullptr;
    es.push_back(e);
  }
  GmgArgs<T> a;
  int nwg;
  if (Stage(es, false, &a, &nwg) != PTH_COMPLETE) return PTH_FAIL;
  const int ne = a.ne;
  {
    ProfScope prof("gmg_outer", stream);
    hipLaunchKernelGGL(OuterKernel<T>, dim3(nwg), dim3(kWG), 0, stream, a);
  }
  hipLaunchKernelGGL(NormKernel<T>, dim3(ne), dim3(kNormWG), 0, stream, a, norms_dev, kWG);
  PTH_HIP(hipGetLastError());
  PTH_HIP(hipMemcpyAsync(norms_host, norms_dev, sizeof(double) * ne, hipMemcpyDeviceToHost,
                         stream));
  PTH_HIP(hipStreamSynchronize(stream));
  stage_used = 0;  // every queued copy has run
  for (size_t q = 0; q < ps.size(); ++q) {
    if (correct[q]) ps[q]->cur = 1 - ps[q]->cur;
    out[q] = std::sqrt(norms_host[q]);
  }
  return PTH_COMPLETE;
}

// One V-cycle on A e = r from e = 0 for every problem in ps, level by level: the same stage of
// every problem is one launch. The level arrays' x / t roles follow the Jacobi ping-pong.
template <typename T>
pth_status GmgBatch<T>::Cycle(const std::vector<GmgProblem<T> *> &ps, double omega, int pre,
                              int post, bool warm) {
  const T w = static_cast<T>(omega);
  int L = 0;
  for (auto *p : ps) L = std::max(L, p->nbottom);
  auto entries = [&](int l) {
    std::vector<GmgEntry<T>> es;
    for (auto *p : ps)
      if (l < p->nbottom) es.push_back(p->Entry(l));
    return es;
  };
  auto sweep = [&](int l) -> pth_status {  // x -> t, then t becomes x
    GmgArgs<T> a;
    int nwg;
    const std::vector<GmgEntry<T>> es = entries(l);
    bool zcol = jacobi_columns > 0;  // columns of kZCol planes when every size allows
    for (const auto &e : es) zcol = zcol && e.n % 64 == 0 && e.n % kZCol == 0;
    if (Stage(es, false, &a, &nwg, zcol ? kWG * kZCol : kWG) != PTH_COMPLETE) return PTH_FAIL;
    ProfScope prof(l == 0 ? "gmg_jacobi_l0" : "gmg_jacobi", stream);
    if (zcol)
      hipLaunchKernelGGL((JacobiZKernel<T, kZCol>), dim3(nwg), dim3(kWG), 0, stream, a, w);
    else
      hipLaunchKernelGGL(JacobiKernel<T>, dim3(nwg), dim3(kWG), 0, stream, a, w);
    PTH_HIP(hipGetLastError());
    for (auto *p : ps)
      if (l < p->nbottom) std::swap(p->lv[l].x, p->lv[l].t);
    return PTH_COMPLETE;
  };
  for (int l = 0; l < L; ++l) {
    GmgArgs<T> a;
    int nwg;
    if (warm && l == 0) {
      for (int s = 0; s < pre; ++s)
        if (sweep(l) != PTH_COMPLETE) return PTH_FAIL;
    } else if (pre > 0) {
      const std::vector<GmgEntry<T>> es = entries(l);
      bool zcol = jacobi_columns > 0 && pre >= 2;
      for (const auto &e : es) zcol = zcol && e.n % 64 == 0 && e.n % kZCol == 0;
      if (Stage(es, false, &a, &nwg, zcol ? kWG * kZCol : kWG) != PTH_COMPLETE) return PTH_FAIL;
      ProfScope prof(l == 0 ? "gmg_pre_l0" : "gmg_pre", stream);
      if (zcol)
        hipLaunchKernelGGL((PreZero2ZKernel<T, kZCol>), dim3(nwg), dim3(kWG), 0, stream, a, w);
      else if (pre >= 2)
        hipLaunchKernelGGL((PreZeroKernel<T, true>), dim3(nwg), dim3(kWG), 0, stream, a, w);
      else
        hipLaunchKernelGGL((PreZeroKernel<T, false>), dim3(nwg), dim3(kWG), 0, stream, a, w);
      PTH_HIP(hipGetLastError());
    } else {
      if (Stage(entries(l), false, &a, &nwg) != PTH_COMPLETE) return PTH_FAIL;
      hipLaunchKernelGGL(ZeroKernel<T>, dim3(nwg), dim3(kWG), 0, stream, a);
      PTH_HIP(hipGetLastError());
    }
    if (!(warm && l == 0))
      for (int s = 2; s < pre; ++s)
        if (sweep(l) != PTH_COMPLETE) return PTH_FAIL;
    if (Stage(entries(l), true, &a, &nwg) != PTH_COMPLETE) return PTH_FAIL;
    ProfScope prof(l == 0 ? "gmg_restrict_l0" : "gmg_restrict", stream);
    hipLaunchKernelGGL(ResidualRestrictKernel<T>, dim3(nwg), dim3(kWG), 0, stream, a);
    PTH_HIP(hipGetLastError());
  }
  {
    std::vector<GmgEntry<T>> es;
    for (auto *p : ps) {
      es.push_back(p->Entry(p->nbottom));
      es.back().warm = warm && p->nbottom == 0;
    }
    GmgArgs<T> a;
    int nwg;
    if (Stage(es, false, &a, &nwg) != PTH_COMPLETE) return PTH_FAIL;
    ProfScope prof("gmg_bottom", stream);
    hipLaunchKernelGGL((BottomKernel<T, GmgBottom<T>::kN>), dim3(a.ne), dim3(1024), 0, stream,
                       a, w, pre, post);
    PTH_HIP(hipGetLastError());
  }
  for (int l = L - 1; l >= 0; --l) {
    GmgArgs<T> a;
    int nwg;
    const std::vector<GmgEntry<T>> es = entries(l);
    // prolongation inside the first sweep (FP32 levels: in FP64 the recomputed neighbours'
    // prolongation costs more than the separate pass saves, measured 208 vs 187 us at 256^3)
    bool fuse = sizeof(T) == 4 && jacobi_columns > 0 && post >= 1;
    for (const auto &e : es) fuse = fuse && e.n % 64 == 0 && e.n % kPJCol == 0;
    if (fuse) {
      if (Stage(es, false, &a, &nwg, kWG * kPJCol) != PTH_COMPLETE) return PTH_FAIL;
      {
        ProfScope prof(l == 0 ? "gmg_prolong_jacobi_l0" : "gmg_prolong_jacobi", stream);
        hipLaunchKernelGGL((ProlongJacobiZKernel<T, kPJCol>), dim3(nwg), dim3(kWG), 0, stream,
                           a, w);
        PTH_HIP(hipGetLastError());
      }
      for (auto *p : ps)
        if (l < p->nbottom) std::swap(p->lv[l].x, p->lv[l].t);
    } else {
      if (Stage(es, false, &a, &nwg) != PTH_COMPLETE) return PTH_FAIL;
      ProfScope prof(l == 0 ? "gmg_prolong_l0" : "gmg_prolong", stream);
      hipLaunchKernelGGL(ProlongKernel<T>, dim3(nwg), dim3(kWG), 0, stream, a);
      PTH_HIP(hipGetLastError());
    }
    for (int s = fuse ? 1 : 0; s < post; ++s)
      if (sweep(l) != PTH_COMPLETE) return PTH_FAIL;
  }
  return PTH_COMPLETE;
}

template <typename T>
pth_status GmgBatch<T>::SetBlobs(GmgProblem<T> &p, std::uint64_t seed) {
  BlobParams P;
  GmgBlobs(seed, P.v);
  hipLaunchKernelGGL(BlobKernel, dim3(Blocks(static_cast<std::int64_t>(p.N) * p.N * p.N)),
                     dim3(kWG), 0, stream, p.b64, p.N, P);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

// u = 0 on the finest level (the solve's initial guess)
template <typename T>
pth_status GmgBatch<T>::Reset(GmgProblem<T> &p) {
  const std::int64_t c = static_cast<std::int64_t>(p.N) * p.N * p.N;
  hipLaunchKernelGGL(FillKernel, dim3(Blocks(c)), dim3(kWG), 0, stream, p.u[0], c);
  PTH_HIP(hipGetLastError());
  p.cur = 0;
  return PTH_COMPLETE;
}

template struct GmgProblem<float>;
template struct GmgProblem<double>;
template struct GmgBatch<float>;
template struct GmgBatch<double>;

}  // namespace pth
