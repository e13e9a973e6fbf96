// TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference's geometric-multigrid Poisson
// solver used as the parity checker of the GPU solver. Never linked into the product library.
//
// Restated from (paths relative to the reference root):
//   operator  example/poisson_gmg/poisson_equation.hpp:131-205 (CalculateFluxes, then
//             FluxMultiplyMatrix; D = 1 on every face, diagonal_alpha = 0), with the correct
//             X3 flux at the top face (the reference writes it into the X2 slot, :153-157)
//   diagonal  poisson_equation.hpp:86-125 (SetDiagonal)
//   Jacobi    src/solvers/mg_solver.hpp:222-297 (weighted, robust::ratio, robust.hpp:59-62)
//   V-cycle   mg_solver.hpp:375-520: pre-smooth, residual, restrict, coarser level,
//             prolongate and correct, post-smooth; the coarsest level only smooths. The
//             reference cycles the full solution (FAS); for this linear operator the same
//             cycle on the residual equation (correction scheme, used here) gives the same
//             iterates in exact arithmetic.
//   transfer  RestrictAverage (src/prolong_restrict/pr_ops.hpp:105-164) for the residual,
//             ProlongateSharedLinear (pr_ops.hpp:166-289) for the correction
//   boundary  FixedFace value 0 (src/bvals/boundary_conditions_generic.hpp:174-255):
//             ghost = 2 * 0 - mirror cell, on every level
//   hierarchy src/mesh/mesh-gmg.cpp:103-154 with the forest of forest.cpp:66-105: on a
//             uniform N^3 mesh every level halves the cells per dimension while the count
//             stays even (256 -> ... -> 1; 192 -> ... -> 3), whatever the block size.
// Smoother weight and sweep counts are parameters (the reference fixes SRJ weights,
// mg_solver.hpp:312-345); the stop test is the relative L2 residual ||b - Ax|| / ||b||.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace orc {
namespace gmg {

struct Level {
  int n = 0;       // cells per dimension
  double h = 0.0;  // cell width on [0, 1]^3
  std::vector<double> x, b, r, t;  // (n+2)^3 arrays with one ghost layer
  std::int64_t Idx(int k, int j, int i) const {
    return ((std::int64_t)(k + 1) * (n + 2) + (j + 1)) * (n + 2) + (i + 1);
  }
  std::int64_t Size() const { return (std::int64_t)(n + 2) * (n + 2) * (n + 2); }
};

// FixedFace 0 on all six faces: ghost = -mirror (edges and corners are not read by the
// 7-point operator or by the face-neighbour prolongation stencil).
inline void ApplyBC(Level &L, std::vector<double> &a) {
  const int n = L.n;
  for (int p = 0; p < n; ++p)
    for (int q = 0; q < n; ++q) {
      a[L.Idx(p, q, -1)] = 2.0 * 0.0 - a[L.Idx(p, q, 0)];
      a[L.Idx(p, q, n)] = 2.0 * 0.0 - a[L.Idx(p, q, n - 1)];
      a[L.Idx(p, -1, q)] = 2.0 * 0.0 - a[L.Idx(p, 0, q)];
      a[L.Idx(p, n, q)] = 2.0 * 0.0 - a[L.Idx(p, n - 1, q)];
      a[L.Idx(-1, p, q)] = 2.0 * 0.0 - a[L.Idx(0, p, q)];
      a[L.Idx(n, p, q)] = 2.0 * 0.0 - a[L.Idx(n - 1, p, q)];
    }
}

// A x at one interior cell in the reference's flux form and operation order (D = 1).
inline double ApplyA(const Level &L, const std::vector<double> &x, int k, int j, int i) {
  const double D = 1.0, alpha = 0.0, dx = L.h;
  const double u = x[L.Idx(k, j, i)];
  double out = -alpha * u;
  const double f1l = D / dx * (x[L.Idx(k, j, i - 1)] - u);
  const double f1r = D / dx * (u - x[L.Idx(k, j, i + 1)]);
  out += (f1l - f1r) / dx;
  const double f2l = D * (x[L.Idx(k, j - 1, i)] - u) / dx;
  const double f2r = D * (u - x[L.Idx(k, j + 1, i)]) / dx;
  out += (f2l - f2r) / dx;
  const double f3l = D * (x[L.Idx(k - 1, j, i)] - u) / dx;
  const double f3r = D * (u - x[L.Idx(k + 1, j, i)]) / dx;
  out += (f3l - f3r) / dx;
  return out;
}

inline double Diagonal(const Level &L) {  // SetDiagonal with D = 1, alpha = 0
  const double D = 1.0, alpha = 0.0, dx = L.h;
  double d = -(D + D) / (dx * dx) - alpha;
  d -= (D + D) / (dx * dx);
  d -= (D + D) / (dx * dx);
  return d;
}

inline double Ratio(double a, double b) {  // robust::ratio
  const double sgn = b >= 0 ? 1.0 : -1.0;
  return a / (b + sgn * (10 * std::numeric_limits<double>::min()));
}

class Solver {
 public:
  std::vector<Level> lv;
  double omega = 2.0 / 3.0;
  int pre = 2, post = 2;
  std::vector<double> history;  // relative residual before each V-cycle, then the final one

  explicit Solver(int N) {
    int n = N;
    while (true) {
      Level L;
      L.n = n;
      L.h = 1.0 / n;
      const std::int64_t s = L.Size();
      L.x.assign(s, 0.0);
      L.b.assign(s, 0.0);
      L.r.assign(s, 0.0);
      L.t.assign(s, 0.0);
      lv.push_back(std::move(L));
      if (n % 2 != 0 || n == 1) break;
      n /= 2;
    }
  }

  int N() const { return lv[0].n; }

  // Cell-centre RHS: sum over 16 Gaussian blobs s exp(-|x - c|^2 / w^2), c ~ U(0.1, 0.9)^3,
  // w ~ U(0.02, 0.08), s = -1 if U < 0.5 else +1, drawn from SplitMix64(seed) in blob order
  // (cx, cy, cz, w, s).
  static void Blobs(std::uint64_t seed, double c[16][3], double w[16], double s[16]) {
    std::uint64_t st = seed;
    auto next = [&]() {
      std::uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      return static_cast<double>((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
    };
    for (int m = 0; m < 16; ++m) {
      for (int d = 0; d < 3; ++d) c[m][d] = 0.1 + 0.8 * next();
      w[m] = 0.02 + 0.06 * next();
      s[m] = next() < 0.5 ? -1.0 : 1.0;
    }
  }

  void SetRhsBlobs(std::uint64_t seed) {
    double c[16][3], w[16], s[16];
    Blobs(seed, c, w, s);
    Level &L = lv[0];
    const int n = L.n;
#pragma omp parallel for
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double x = (i + 0.5) * L.h, y = (j + 0.5) * L.h, z = (k + 0.5) * L.h;
          double v = 0.0;
          for (int m = 0; m < 16; ++m) {
            const double dx = x - c[m][0], dy = y - c[m][1], dz = z - c[m][2];
            v += s[m] * std::exp(-(dx * dx + dy * dy + dz * dz) / (w[m] * w[m]));
          }
          L.b[L.Idx(k, j, i)] = v;
        }
  }

  // out = A x on the interior of level l (ghosts of x are refreshed first)
  void Apply(int l, std::vector<double> &x, std::vector<double> &out) {
    Level &L = lv[l];
    ApplyBC(L, x);
    const int n = L.n;
#pragma omp parallel for
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) out[L.Idx(k, j, i)] = ApplyA(L, x, k, j, i);
  }

  // one weighted Jacobi sweep on level l: x <- w (b - (Ax - D x)) / D + (1 - w) x
  void Jacobi(int l) {
    Level &L = lv[l];
    ApplyBC(L, L.x);
    const double D = Diagonal(L), w = omega;
    const int n = L.n;
#pragma omp parallel for
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const std::int64_t c = L.Idx(k, j, i);
          const double xo = L.x[c];
          const double off = ApplyA(L, L.x, k, j, i) - D * xo;
          const double val = L.b[c] - off;
          L.t[c] = w * Ratio(val, D) + (1.0 - w) * xo;
        }
    std::swap(L.x, L.t);
  }

  // r = b - A x on level l
  void Residual(int l) {
    Level &L = lv[l];
    ApplyBC(L, L.x);
    const int n = L.n;
#pragma omp parallel for
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const std::int64_t c = L.Idx(k, j, i);
          L.r[c] = L.b[c] - ApplyA(L, L.x, k, j, i);
        }
  }

  // coarse b = RestrictAverage(fine r)
  void Restrict(int l) {
    Level &F = lv[l], &C = lv[l + 1];
    const double vol = F.h * F.h * F.h;
    const int n = C.n;
#pragma omp parallel for
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double v[2][2][2], t[2][2][2];
          for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb)
              for (int cc = 0; cc < 2; ++cc) {
                v[a][bb][cc] = vol;
                t[a][bb][cc] = vol * F.r[F.Idx(2 * k + a, 2 * j + bb, 2 * i + cc)];
              }
          const double tv = ((v[0][0][0] + v[0][1][0]) + (v[0][0][1] + v[0][1][1])) +
                            ((v[1][0][0] + v[1][1][0]) + (v[1][0][1] + v[1][1][1]));
          C.b[C.Idx(k, j, i)] = (((t[0][0][0] + t[0][1][0]) + (t[0][0][1] + t[0][1][1])) +
                                 ((t[1][0][0] + t[1][1][0]) + (t[1][0][1] + t[1][1][1]))) /
                                tv;
        }
  }

  // fine x += ProlongateSharedLinear(coarse x); coarse ghosts from the boundary condition
  void ProlongAdd(int l) {
    Level &F = lv[l], &C = lv[l + 1];
    ApplyBC(C, C.x);
    const double hc = C.h, df = 0.25 * C.h;  // centre spacing and child offset
    const int n = C.n;
#pragma omp parallel for
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double fc = C.x[C.Idx(k, j, i)];
          const double g1m = (fc - C.x[C.Idx(k, j, i - 1)]) / hc;
          const double g1p = (C.x[C.Idx(k, j, i + 1)] - fc) / hc;
          const double g2m = (fc - C.x[C.Idx(k, j - 1, i)]) / hc;
          const double g2p = (C.x[C.Idx(k, j + 1, i)] - fc) / hc;
          const double g3m = (fc - C.x[C.Idx(k - 1, j, i)]) / hc;
          const double g3p = (C.x[C.Idx(k + 1, j, i)] - fc) / hc;
          double e[2][2][2];
          e[0][0][0] = fc - (g1m * df + g2m * df + g3m * df);
          e[0][0][1] = fc + (g1p * df - g2m * df - g3m * df);
          e[0][1][0] = fc - (g1m * df - g2p * df + g3m * df);
          e[0][1][1] = fc + (g1p * df + g2p * df - g3m * df);
          e[1][0][0] = fc - (g1m * df + g2m * df - g3p * df);
          e[1][0][1] = fc + (g1p * df - g2m * df + g3p * df);
          e[1][1][0] = fc - (g1m * df - g2p * df - g3p * df);
          e[1][1][1] = fc + (g1p * df + g2p * df + g3p * df);
          for (int a = 0; a < 2; ++a)
            for (int bb = 0; bb < 2; ++bb)
              for (int cc = 0; cc < 2; ++cc)
                F.x[F.Idx(2 * k + a, 2 * j + bb, 2 * i + cc)] += e[a][bb][cc];
        }
  }

  // approximate solve of A x = b on level l from x = 0
  void VCycle(int l) {
    Level &L = lv[l];
    std::fill(L.x.begin(), L.x.end(), 0.0);
    for (int s = 0; s < pre; ++s) Jacobi(l);
    if (l + 1 == (int)lv.size()) return;  // coarsest: smoothing only
    Residual(l);
    Restrict(l);
    VCycle(l + 1);
    ProlongAdd(l);
    for (int s = 0; s < post; ++s) Jacobi(l);
  }

  static double Norm2(const Level &L, const std::vector<double> &a) {
    double acc = 0.0;
    const int n = L.n;
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          const double v = a[L.Idx(k, j, i)];
          acc += v * v;
        }
    return std::sqrt(acc);
  }

  // Outer iteration u <- u + V(b - A u) until ||b - A u|| / ||b|| < tol; u, bf are the
  // finest-level solution and right-hand side (kept apart from the level arrays that the
  // cycle reuses). Returns the number of V-cycles.
  int Solve(std::vector<double> &u, const std::vector<double> &bf, double tol, int maxit) {
    Level &F = lv[0];
    history.clear();
    const double bn = Norm2(F, bf);
    std::vector<double> Au(F.Size(), 0.0);
    int it = 0;
    for (;; ++it) {
      Apply(0, u, Au);
      const int n = F.n;
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const std::int64_t c = F.Idx(k, j, i);
            F.b[c] = bf[c] - Au[c];
          }
      const double rel = Norm2(F, F.b) / bn;
      history.push_back(rel);
      if (rel < tol || it >= maxit) break;
      VCycle(0);
      for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
          for (int i = 0; i < n; ++i) {
            const std::int64_t c = F.Idx(k, j, i);
            u[c] += F.x[c];
          }
    }
    return it;
  }
};

}  // namespace gmg
}  // namespace orc
