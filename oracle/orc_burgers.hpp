// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU restatement of the Burgers benchmark (Parthenon-VIBE) stage kernels and its RK2
// task ordering, over the mesh and exchange restatements in orc_mesh.hpp / orc_bvals.hpp.
//
// Restated from (paths relative to the reference root):
//   benchmarks/burgers/recon.hpp:28-99            mc, Linear, WENO5Z
//   benchmarks/burgers/burgers_package.hpp:33-43  lr_to_flux (HLL)
//   benchmarks/burgers/burgers_package.cpp:143-167  CalculateDerived
//   benchmarks/burgers/burgers_package.cpp:170-200  EstimateTimestepMesh
//   benchmarks/burgers/burgers_package.cpp:202-404  CalculateFluxes
//   benchmarks/burgers/burgers_package.cpp:406-439  MassHistory
//   benchmarks/burgers/burgers_driver.cpp:53-148  stage task order
//   src/interface/update.hpp:43-57,71-90,122-137  FluxDivHelper, WeightedSumData
//   src/interface/update.cpp:63-84                FluxDivergence<MeshData>
//   src/time_integration/low_storage_integrator.cpp:67-83  rk2 beta = {1, 1/2}
//   src/driver/driver.cpp:210-235                 SetGlobalTimeStep (dt growth <= 2x)
//   src/utils/robust.hpp:38-41                    EPS = 10 * DBL_EPSILON
#pragma once
#include <cfloat>
#include <limits>

#include "orc_bvals.hpp"

namespace orc {

constexpr double kEPS = 10.0 * DBL_EPSILON;  // robust::EPS() = 10 * numeric_limits::epsilon()

inline double mc(double dm, double dp) {
  const double dc = (dm * dp > 0.0) * 0.5 * (dm + dp);
  return std::copysign(std::min(std::fabs(dc), 2.0 * std::min(std::fabs(dm), std::fabs(dp))),
                       dc);
}

inline void Linear(double qm, double q0, double qp, double &ql, double &qr) {
  double dq = qp - q0;
  dq = 0.5 * mc(q0 - qm, dq);
  ql = q0 + dq;
  qr = q0 - dq;
}

inline void WENO5Z(double q0, double q1, double q2, double q3, double q4, double &ql,
                   double &qr) {
  const double a00 = 1.0 / 3.0, a01 = -7.0 / 6.0, a02 = 11.0 / 6.0;
  const double a10 = -1.0 / 6.0, a11 = 5.0 / 6.0, a12 = 1.0 / 3.0;
  const double a20 = 1.0 / 3.0, a21 = 5.0 / 6.0, a22 = -1.0 / 6.0;
  const double g0 = 0.1, g1 = 0.6, g2 = 0.3;
  const double eps = kEPS;
  const double t13 = 13.0 / 3.0;
  double a = q0 - 2 * q1 + q2;
  double b = q0 - 4.0 * q1 + 3.0 * q2;
  double beta0 = t13 * a * a + b * b + eps;
  a = q1 - 2.0 * q2 + q3;
  b = q3 - q1;
  double beta1 = t13 * a * a + b * b + eps;
  a = q2 - 2.0 * q3 + q4;
  b = q4 - 4.0 * q3 + 3.0 * q2;
  double beta2 = t13 * a * a + b * b + eps;
  const double tau5 = std::abs(beta2 - beta0);
  beta0 = (beta0 + tau5) / beta0;
  beta1 = (beta1 + tau5) / beta1;
  beta2 = (beta2 + tau5) / beta2;
  double w0 = g0 * beta0 + eps;
  double w1 = g1 * beta1 + eps;
  double w2 = g2 * beta2 + eps;
  double wsum = 1.0 / (w0 + w1 + w2);
  ql = w0 * (a00 * q0 + a01 * q1 + a02 * q2);
  ql += w1 * (a10 * q1 + a11 * q2 + a12 * q3);
  ql += w2 * (a20 * q2 + a21 * q3 + a22 * q4);
  ql *= wsum;
  const double alpha_l = 3.0 * wsum * w0 * w1 * w2 / (g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;
  w0 = g0 * beta2 + eps;
  w1 = g1 * beta1 + eps;
  w2 = g2 * beta0 + eps;
  wsum = 1.0 / (w0 + w1 + w2);
  qr = w0 * (a00 * q4 + a01 * q3 + a02 * q2);
  qr += w1 * (a10 * q3 + a11 * q2 + a12 * q1);
  qr += w2 * (a20 * q2 + a21 * q1 + a22 * q0);
  qr *= wsum;
  const double alpha_r = 3.0 * wsum * w0 * w1 * w2 / (g2 * w0 * w1 + g1 * w0 * w2 + g0 * w1 * w2) + eps;
  double dq = q3 - q2;
  dq = 0.5 * mc(q2 - q1, dq);
  const double alpha_lin = 2.0 * alpha_l * alpha_r / (alpha_l + alpha_r);
  ql = alpha_lin * ql + (1.0 - alpha_lin) * (q2 + dq);
  qr = alpha_lin * qr + (1.0 - alpha_lin) * (q2 - dq);
}

inline void lr_to_flux(double uxl, double uxr, double uyl, double uyr, double uzl, double uzr,
                       double upl, double upr, double &sl, double &sr, double &fux, double &fuy,
                       double &fuz) {
  sl = std::min(std::min(upl, upr), 0.0);
  sr = std::max(std::max(upl, upr), 0.0);
  const double islsr = 1.0 / (sr - sl + (sl * sr == 0.0));
  fux = 0.5 * (sr * uxl * upl - sl * uxr * upr + sl * sr * (uxr - uxl)) * islsr;
  fuy = 0.5 * (sr * uyl * upl - sl * uyr * upr + sl * sr * (uyr - uyl)) * islsr;
  fuz = 0.5 * (sr * uzl * upl - sl * uzr * upr + sl * sr * (uzr - uzl)) * islsr;
}

// The scalar (passive) flux of burgers_package.cpp:335-340 with the face's sl, sr.
inline double qflux(double upl, double upr, double ql, double qr, double sl, double sr) {
  return (sr * upl * ql - sl * upr * qr + sl * sr * (qr - ql)) / (sr - sl + (sl * sr == 0.0));
}

// State of one Burgers container ("base", "1") : U and its three face-flux arrays.
struct Container {
  Field U;
  std::vector<double> flux[3];   // same brick shape as U, face i stored at index i
  std::vector<double> cflux[3];  // coarse buffer of the flux variable (multilevel)
};

struct Burgers {
  const Mesh *m;
  int V;
  double cfl;
  int recon;  // 0 = weno5, 1 = linear
  Container base, one, dudt;
  std::vector<double> face[6];  // Ulx, Urx, Uly, Ury, Ulz, Urz (OneCopy scratch)
  std::vector<double> derived;  // OneCopy
  Exchange ex_base, ex_one;
  FluxCorrection flxcor;
  bool flux_correction = true;  // the flxcor tasks of burgers_driver.cpp:84-96 (multilevel)
  std::vector<Field> fb, fo;    // field lists for the exchanges
  double dt = std::numeric_limits<double>::max(), time = 0.0;
  int ncycle = 0;
  // <parthenon/time> limits with the defaults of EvolutionDriver (src/driver/driver.hpp:66-94)
  struct Limits {
    double tlim = std::numeric_limits<double>::infinity();
    double dt_init = std::numeric_limits<double>::max();
    bool dt_init_force = false;
    double dt_user = std::numeric_limits<double>::max();
    double dt_force = std::numeric_limits<double>::lowest();
    double dt_floor = std::numeric_limits<double>::min();
    double dt_ceil = std::numeric_limits<double>::max();
    int nlim = -1;
  } lim;

  std::int64_t slab() const { return m->cellbounds.size(); }

  void Init(const Mesh &mesh, int num_scalars, double cfl_, int recon_) {
    m = &mesh;
    V = 3 + num_scalars;
    cfl = cfl_;
    recon = recon_;
    const std::int64_t nb = mesh.blocks.size();
    for (Container *c : {&base, &one, &dudt}) {
      c->U.name = "U";
      c->U.ncomp = V;
      c->U.data.assign(nb * V * slab(), 0.0);
      if (mesh.multilevel) c->U.coarse.assign(nb * V * mesh.c_cellbounds.size(), 0.0);
      for (int d = 0; d < 3; ++d) {
        c->flux[d].assign(nb * V * slab(), 0.0);
        if (mesh.multilevel) c->cflux[d].assign(nb * V * mesh.c_cellbounds.size(), 0.0);
      }
    }
    for (int q = 0; q < 6; ++q) face[q].assign(nb * V * slab(), 0.0);
    derived.assign(nb * slab(), 0.0);
  }

  double &u(std::vector<double> &a, int b, int n, int k, int j, int i) const {
    const Shape &s = m->cellbounds;
    return a[((((std::int64_t)b * V + n) * s.ent(2) + k) * s.ent(1) + j) * s.ent(0) + i];
  }

  // burgers_package.cpp:202-404
  void CalculateFluxes(Container &c) {
    const Shape &s = m->cellbounds;
    const int ndim = m->ndim;
    const int ibs = s.s(0), ibe = s.e(0), jbs = s.s(1), jbe = s.e(1), kbs = s.s(2), kbe = s.e(2);
    const int dk = ndim > 2 ? 1 : 0, dj = ndim > 1 ? 1 : 0;
    auto &U = c.U.data;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < (int)m->blocks.size(); ++b)
      for (int k = kbs - dk; k <= kbe + dk; ++k)
        for (int j = jbs - dj; j <= jbe + dj; ++j) {
          const bool xrec = (k >= kbs && k <= kbe) && (j >= jbs && j <= jbe);
          const bool yrec = (k >= kbs && k <= kbe) && (ndim > 1);
          const bool zrec = (j >= jbs && j <= jbe) && (ndim > 2);
          for (int n = 0; n < V; ++n) {
            if (xrec)
              for (int i = ibs - 1; i <= ibe + 1; ++i) {
                double ql, qr;
                if (recon == 0)
                  WENO5Z(u(U, b, n, k, j, i - 2), u(U, b, n, k, j, i - 1), u(U, b, n, k, j, i),
                         u(U, b, n, k, j, i + 1), u(U, b, n, k, j, i + 2), ql, qr);
                else
                  Linear(u(U, b, n, k, j, i - 1), u(U, b, n, k, j, i), u(U, b, n, k, j, i + 1), ql,
                         qr);
                u(face[0], b, n, k, j, i + 1) = ql;
                u(face[1], b, n, k, j, i) = qr;
              }
            if (yrec)
              for (int i = ibs; i <= ibe; ++i) {
                double ql, qr;
                if (recon == 0)
                  WENO5Z(u(U, b, n, k, j - 2, i), u(U, b, n, k, j - 1, i), u(U, b, n, k, j, i),
                         u(U, b, n, k, j + 1, i), u(U, b, n, k, j + 2, i), ql, qr);
                else
                  Linear(u(U, b, n, k, j - 1, i), u(U, b, n, k, j, i), u(U, b, n, k, j + 1, i), ql,
                         qr);
                u(face[2], b, n, k, j + 1, i) = ql;
                u(face[3], b, n, k, j, i) = qr;
              }
            if (zrec)
              for (int i = ibs; i <= ibe; ++i) {
                double ql, qr;
                if (recon == 0)
                  WENO5Z(u(U, b, n, k - 2, j, i), u(U, b, n, k - 1, j, i), u(U, b, n, k, j, i),
                         u(U, b, n, k + 1, j, i), u(U, b, n, k + 2, j, i), ql, qr);
                else
                  Linear(u(U, b, n, k - 1, j, i), u(U, b, n, k, j, i), u(U, b, n, k + 1, j, i), ql,
                         qr);
                u(face[4], b, n, k + 1, j, i) = ql;
                u(face[5], b, n, k, j, i) = qr;
              }
          }
        }
    // HLL fluxes; per-face sl, sr reused by the scalar fluxes (burgers_package.cpp:307-401)
#pragma omp parallel for schedule(static)
    for (int b = 0; b < (int)m->blocks.size(); ++b)
      for (int k = kbs; k <= kbe + dk; ++k)
        for (int j = jbs; j <= jbe + dj; ++j) {
          const bool xflux = (k <= kbe && j <= jbe);
          const bool yflux = (ndim > 1 && k <= kbe);
          const bool zflux = (ndim > 2 && j <= jbe);
          for (int d = 0; d < 3; ++d) {
            const bool doit = d == 0 ? xflux : (d == 1 ? yflux : zflux);
            if (!doit) continue;
            auto &L = face[2 * d];
            auto &R = face[2 * d + 1];
            const int ie = d == 0 ? ibe + 1 : ibe;
            for (int i = ibs; i <= ie; ++i) {
              double sl, sr, f0, f1, f2;
              lr_to_flux(u(L, b, 0, k, j, i), u(R, b, 0, k, j, i), u(L, b, 1, k, j, i),
                         u(R, b, 1, k, j, i), u(L, b, 2, k, j, i), u(R, b, 2, k, j, i),
                         u(L, b, d, k, j, i), u(R, b, d, k, j, i), sl, sr, f0, f1, f2);
              u(c.flux[d], b, 0, k, j, i) = f0;
              u(c.flux[d], b, 1, k, j, i) = f1;
              u(c.flux[d], b, 2, k, j, i) = f2;
              const double upl = u(L, b, d, k, j, i), upr = u(R, b, d, k, j, i);
              for (int n = 3; n < V; ++n)
                u(c.flux[d], b, n, k, j, i) =
                    qflux(upl, upr, u(L, b, n, k, j, i), u(R, b, n, k, j, i), sl, sr);
            }
          }
        }
  }

  // update.cpp:63-84 with update.hpp:43-57
  void FluxDivergence(Container &in, Container &out) {
    const Shape &s = m->cellbounds;
    const int ndim = m->ndim;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < (int)m->blocks.size(); ++b) {
      const Coords &co = m->blocks[b].coords;
      for (int n = 0; n < V; ++n)
        for (int k = s.s(2); k <= s.e(2); ++k)
          for (int j = s.s(1); j <= s.e(1); ++j)
            for (int i = s.s(0); i <= s.e(0); ++i) {
              double du = (co.area[0] * u(in.flux[0], b, n, k, j, i + 1) -
                           co.area[0] * u(in.flux[0], b, n, k, j, i));
              if (ndim >= 2)
                du += (co.area[1] * u(in.flux[1], b, n, k, j + 1, i) -
                       co.area[1] * u(in.flux[1], b, n, k, j, i));
              if (ndim == 3)
                du += (co.area[2] * u(in.flux[2], b, n, k + 1, j, i) -
                       co.area[2] * u(in.flux[2], b, n, k, j, i));
              u(out.U.data, b, n, k, j, i) = -du / co.vol;
            }
    }
  }

  // update.hpp:71-90 over entire arrays of the Independent variable U
  static void WeightedSum(const std::vector<double> &x, const std::vector<double> &y, double w1,
                          double w2, std::vector<double> &z) {
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < z.size(); ++q) z[q] = w1 * x[q] + w2 * y[q];
  }

  // burgers_package.cpp:143-167 (pack order {"derived","U"}: v(b,4) is U[3])
  void CalculateDerived(Container &c) {
    const Shape &s = m->cellbounds;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < (int)m->blocks.size(); ++b)
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i) {
            const double u1 = u(c.U.data, b, 0, k, j, i), u2 = u(c.U.data, b, 1, k, j, i),
                         u3 = u(c.U.data, b, 2, k, j, i), d = u(c.U.data, b, 3, k, j, i);
            derived[(((std::int64_t)b * s.ent(2) + k) * s.ent(1) + j) * s.ent(0) + i] =
                0.5 * d * (u1 * u1 + u2 * u2 + u3 * u3);
          }
  }

  // burgers_package.cpp:170-200
  double EstimateTimestep(Container &c) {
    const Shape &s = m->cellbounds;
    const int ndim = m->ndim;
    double mn = std::numeric_limits<double>::max();
#pragma omp parallel for schedule(static) reduction(min : mn)
    for (int b = 0; b < (int)m->blocks.size(); ++b) {
      const Coords &co = m->blocks[b].coords;
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i)
            mn = std::min(mn, 1.0 / ((std::abs(u(c.U.data, b, 0, k, j, i))) / co.dx[0] +
                                     (ndim > 1) * (std::abs(u(c.U.data, b, 1, k, j, i))) / co.dx[1] +
                                     (ndim > 2) * (std::abs(u(c.U.data, b, 2, k, j, i))) / co.dx[2]));
    }
    return cfl * mn;
  }

  void DoExchange(Container &c, Exchange &ex, std::vector<Field> &flist) {
    flist.resize(1);
    std::swap(flist[0], c.U);
    ex.WithBoundaryConditions(flist);
    std::swap(flist[0], c.U);
  }

  void BuildExchanges() {
    fb.resize(1);
    std::swap(fb[0], base.U);
    ex_base.Build(*m, fb);
    std::swap(fb[0], base.U);
    fo.resize(1);
    std::swap(fo[0], one.U);
    ex_one.Build(*m, fo);
    std::swap(fo[0], one.U);
    flxcor.Build(*m, V);
  }

  // Mesh::Initialize after the problem generator: exchange, FillDerived, first dt
  void Initialize() {
    BuildExchanges();
    DoExchange(base, ex_base, fb);
    CalculateDerived(base);
    dt = std::numeric_limits<double>::max();
    SetGlobalTimeStep(EstimateTimestep(base));
  }

  // EvolutionDriver::SetGlobalTimeStep (driver.cpp:210-269) with dt_factor 2; the dt_min /
  // dt_max cycle-limit checks only abort and change no value.
  void SetGlobalTimeStep(double block_dt) {
    if (lim.dt_force > 0.0) {
      dt = lim.dt_force;
    } else if (ncycle == 0 && lim.dt_init_force && lim.dt_init > 0.0) {
      dt = lim.dt_init;
    } else {
      if (dt < 0.1 * std::numeric_limits<double>::max()) dt *= 2.0;
      if (ncycle == 0) dt = std::min(dt, lim.dt_init);
      dt = std::min(dt, block_dt);
      dt = std::min(dt, lim.dt_user);
      dt = std::max(lim.dt_floor, std::min(dt, lim.dt_ceil));
    }
    if (time < lim.tlim && (lim.tlim - time) < dt) dt = lim.tlim - time;
  }

  // SimTime::KeepGoing (src/basic_types.hpp:241): time < tlim and ncycle < nlim
  bool KeepGoing() const { return time < lim.tlim && (lim.nlim < 0 || ncycle < lim.nlim); }

  // One RK2 cycle in the task order of burgers_driver.cpp:53-148: fluxes, flux correction
  // (multilevel), divergence, averaging, update, exchange (+ restriction / prolongation),
  // derived field, time-step estimate on the last stage.
  void Cycle() {
    double new_dt = 0;
    for (int stage = 1; stage <= 2; ++stage) {
      const double beta = stage == 1 ? 1.0 : 0.5;
      Container &mc0 = stage == 1 ? base : one;
      Container &mc1 = stage == 1 ? one : base;
      Exchange &ex1 = stage == 1 ? ex_one : ex_base;
      std::vector<Field> &fl1 = stage == 1 ? fo : fb;
      CalculateFluxes(mc0);
      if (m->multilevel && flux_correction) flxcor.Run(mc0.flux, mc0.cflux, V);
      FluxDivergence(mc0, dudt);
      WeightedSum(mc0.U.data, base.U.data, beta, 1.0 - beta, mc0.U.data);
      WeightedSum(mc0.U.data, dudt.U.data, 1.0, beta * dt, mc1.U.data);
      DoExchange(mc1, ex1, fl1);
      CalculateDerived(mc1);
      if (stage == 2) new_dt = EstimateTimestep(mc1);
    }
    ncycle++;
    time += dt;
    SetGlobalTimeStep(new_dt);
  }

  // burgers_package.cpp:406-439 (one octant), summed in block/var/k/j/i order
  double MassHistory(const double xmn[3], const double xmx[3], double mesh_vol) {
    const Shape &s = m->cellbounds;
    double result = 0.0;
    for (int b = 0; b < (int)m->blocks.size(); ++b) {
      const Coords &co = m->blocks[b].coords;
      for (int n = 0; n < V; ++n)
        for (int k = s.s(2); k <= s.e(2); ++k)
          for (int j = s.s(1); j <= s.e(1); ++j)
            for (int i = s.s(0); i <= s.e(0); ++i) {
              const double w = co.vol / (mesh_vol + 1e-20);
              const double x1 = co.Xc(0, i), x2 = co.Xc(1, j), x3 = co.Xc(2, k);
              const double mask = (xmn[0] <= x1) && (x1 <= xmx[0]) && (xmn[1] <= x2) &&
                                  (x2 <= xmx[1]) && (xmn[2] <= x3) && (x3 <= xmx[2]);
              const double q = u(base.U.data, b, n, k, j, i);
              result += mask * q * q * w;
            }
    }
    return result;
  }
};

}  // namespace orc
