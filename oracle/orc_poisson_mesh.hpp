// TEST INFRASTRUCTURE ONLY: the Poisson problem of example/poisson_gmg on a block mesh of any
// refinement, as the parity checker of the GPU multigrid solver on refined meshes. Never linked
// into the product library.
//
// Restated from (paths relative to the reference root):
//   leaf operator  the residual check of MGSolver::AddTasks (src/solvers/mg_solver.hpp:
//                  131-141) and the use_exact_rhs set-up of poisson_driver.cpp:87-95: the
//                  boundary exchange of the leaf grid (AddBoundaryExchangeTasks, with u's
//                  registered ProlongateSharedLinear, poisson_package.cpp:122-126, and the
//                  FixedFace-0 conditions, :60-78), then PoissonEquation::Ax =
//                  CalculateFluxes + FluxMultiplyMatrix (poisson_equation.hpp:41-53,
//                  131-205) without flux correction (the deck's flux_correct = false), D
//                  constant or the face field D; the top x3 face flux is the 7-point operator's (the reference
//                  writes it into the x2 slot, :153-157)
//   solve          the fixed point every convergent iteration reaches: A u = rhs on the leaf
//                  grid, by BiCGSTAB (van der Vorst, SIAM J. Sci. Stat. Comput. 13 (1992)
//                  631-644) on the interior cells, which needs no symmetry (linear
//                  prolongation at fine-coarse faces makes the operator non-symmetric)
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "orc_bvals.hpp"

namespace orc {
namespace poisson {

struct LeafProblem {
  const Mesh *m;
  double D = 1.0, alpha = 0.0;
  // the face field D (Metadata::Face, poisson_package.cpp:112-118), or null for the constant
  // D: [block][3][cellbounds], component d at cell (i, j, k) = D on the lower face along d
  const double *Df = nullptr;
  Exchange ex;
  std::vector<Field> fs;  // the work field u with ghosts and coarse buffers (the exchange's
                          // regions point into it, so it stays in place)

  LeafProblem(const Mesh &mesh, double D_, double alpha_)
      : m(&mesh), D(D_), alpha(alpha_), fs(1) {
    Field &u = fs[0];
    u.ncomp = 1;
    u.data.assign(m->blocks.size() * m->cellbounds.size(), 0.0);
    if (m->multilevel) u.coarse.assign(m->blocks.size() * m->c_cellbounds.size(), 0.0);
    ex.Build(*m, fs);
  }

  std::int64_t Cells() const {
    const Shape &s = m->cellbounds;
    return (std::int64_t)m->blocks.size() * s.n[0] * s.n[1] * s.n[2];
  }

  // interior cells of every block, in block then [k][j][i] order <-> the bricks
  void ToBricks(const std::vector<double> &v, std::vector<double> &bricks) const {
    const Shape &s = m->cellbounds;
    std::int64_t q = 0;
    for (const auto &b : m->blocks) {
      double *slab = bricks.data() + (std::int64_t)b.gid * s.size();
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i) At(slab, s, 0, k, j, i) = v[q++];
    }
  }
  void FromBricks(const std::vector<double> &bricks, std::vector<double> &v) const {
    const Shape &s = m->cellbounds;
    v.resize(Cells());
    std::int64_t q = 0;
    for (const auto &b : m->blocks) {
      const double *slab = bricks.data() + (std::int64_t)b.gid * s.size();
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i)
            v[q++] = slab[((std::int64_t)k * s.ent(1) + j) * s.ent(0) + i];
    }
  }

  // out = A x on the interior cells (x: interior values in FromBricks order)
  void Apply(const std::vector<double> &x, std::vector<double> &out) {
    Field &u = fs[0];
    ToBricks(x, u.data);
    ex.WithBoundaryConditions(fs, /*minmod=*/false);
    const Shape &s = m->cellbounds;
    const int ndim = m->ndim;
    const std::int64_t sj = s.ent(0), sk = (std::int64_t)s.ent(0) * s.ent(1);
    out.resize(Cells());
    std::int64_t q = 0;
    for (const auto &b : m->blocks) {
      const double *p = u.data.data() + (std::int64_t)b.gid * s.size();
      const double *df = Df ? Df + (std::int64_t)b.gid * 3 * s.size() : nullptr;
      // D on face (d, cell c): the face field's component d, or the constant
      auto Dat = [&](int d, std::int64_t c) { return df ? df[d * s.size() + c] : D; };
      const double dx1 = b.coords.dx[0], dx2 = b.coords.dx[1], dx3 = b.coords.dx[2];
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i) {
            const std::int64_t c = ((std::int64_t)k * s.ent(1) + j) * s.ent(0) + i;
            double r = -alpha * p[c];
            const double f1l = Dat(0, c) / dx1 * (p[c - 1] - p[c]);
            const double f1r = Dat(0, c + 1) / dx1 * (p[c] - p[c + 1]);
            r += (f1l - f1r) / dx1;
            if (ndim > 1) {
              const double f2l = Dat(1, c) * (p[c - sj] - p[c]) / dx2;
              const double f2r = Dat(1, c + sj) * (p[c] - p[c + sj]) / dx2;
              r += (f2l - f2r) / dx2;
            }
            if (ndim > 2) {
              const double f3l = Dat(2, c) * (p[c - sk] - p[c]) / dx3;
              const double f3r = Dat(2, c + sk) * (p[c] - p[c + sk]) / dx3;
              r += (f3l - f3r) / dx3;
            }
            out[q++] = r;
          }
    }
  }

  static double Dot(const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (size_t q = 0; q < a.size(); ++q) s += a[q] * b[q];
    return s;
  }

  // BiCGSTAB on A e = r from e = 0, until the recurrence residual falls by `drop` or maxit
  // iterations; returns the iterations
  int Inner(const std::vector<double> &r_in, std::vector<double> &e, double drop, int maxit) {
    const std::int64_t n = Cells();
    std::vector<double> r(r_in), r0(r_in), p(n, 0.0), v(n, 0.0), s(n), t(n);
    e.assign(n, 0.0);
    const double stop = drop * std::sqrt(Dot(r, r));
    double rho = 1.0, alph = 1.0, omega = 1.0;
    int it = 0;
    while (it < maxit && std::sqrt(Dot(r, r)) > stop) {
      ++it;
      const double rho1 = Dot(r0, r);
      if (rho1 == 0.0) break;
      const double beta = (rho1 / rho) * (alph / omega);
      for (std::int64_t q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - omega * v[q]);
      Apply(p, v);
      alph = rho1 / Dot(r0, v);
      for (std::int64_t q = 0; q < n; ++q) s[q] = r[q] - alph * v[q];
      Apply(s, t);
      const double tt = Dot(t, t);
      omega = tt > 0.0 ? Dot(t, s) / tt : 0.0;
      for (std::int64_t q = 0; q < n; ++q) {
        e[q] += alph * p[q] + omega * s[q];
        r[q] = s[q] - omega * t[q];
      }
      rho = rho1;
      if (omega == 0.0) break;
    }
    return it;
  }

  // From x (in/out) until the RMS residual sqrt(|b - A x|^2 / cells) <= tol: iterative
  // refinement around BiCGSTAB (the true residual recomputed after every inner solve, which
  // the recurrence residual of one long run drifts away from); returns the BiCGSTAB
  // iterations, or -1 if tol is not reached within maxit of them
  int Solve(const std::vector<double> &b, std::vector<double> &x, double tol, int maxit,
            double *rms_out) {
    const std::int64_t n = Cells();
    std::vector<double> r(n), e, ax;
    const double scale = 1.0 / std::sqrt((double)n);
    int it = 0;
    double rms = 0.0;
    for (;;) {
      Apply(x, ax);
      for (std::int64_t q = 0; q < n; ++q) r[q] = b[q] - ax[q];
      rms = std::sqrt(Dot(r, r)) * scale;
      if (rms <= tol || it >= maxit) break;
      it += Inner(r, e, 1e-6, std::min(400, maxit - it));
      for (std::int64_t q = 0; q < n; ++q) x[q] += e[q];
    }
    if (rms_out) *rms_out = rms;
    return rms <= tol ? it : -1;
  }
};

}  // namespace poisson
}  // namespace orc
