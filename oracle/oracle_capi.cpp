// ORACLE — TEST INFRASTRUCTURE ONLY.
// C entry points over the CPU restatement (orc_*.hpp) so pytest can drive it through ctypes.
// Only tests/, the smoke check and the CPU-baseline leg of bench.py may load liboracle.so;
// the product library never links or calls it.
#include <cstdint>
#include <cstring>
#include <memory>

#include "orc_amr.hpp"
#include "orc_burgers.hpp"
#include "orc_gmg.hpp"
#include "orc_poisson_mesh.hpp"

using namespace orc;

namespace {

// SplitMix64 (Steele, Lea, Flood 2014): the portable seeded generator every input uses.
struct SplitMix64 {
  std::uint64_t s;
  std::uint64_t next() {
    std::uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

struct OMesh {
  std::unique_ptr<Mesh> m;
};

}  // namespace

extern "C" {

void *orc_mesh_new(const int *nx, const int *bnx, int nghost, const double *xmin,
                   const double *xmax, const int *periodic) {
  MeshDesc d;
  for (int k = 0; k < 3; ++k) {
    d.nx[k] = nx[k];
    d.bnx[k] = bnx[k];
    d.xmin[k] = xmin[k];
    d.xmax[k] = xmax[k];
    d.periodic[k] = periodic[k];
    d.bc[2 * k] = d.bc[2 * k + 1] = periodic[k] ? kBcPeriodic : kBcOutflow;
  }
  d.nghost = nghost;
  auto *o = new OMesh;
  o->m.reset(new Mesh(d));
  return o;
}

// Boundary condition of a domain face (0 inner x1, 1 outer x1, ..., 5 outer x3): kBcOutflow or
// kBcReflect on a non-periodic dimension (its faces default to outflow). Returns 0, or -1 if
// the face belongs to a periodic dimension or the type is unknown.
int orc_mesh_set_boundary(void *h, int face, int type) {
  Mesh &m = *static_cast<OMesh *>(h)->m;
  if (face < 0 || face >= 6 || m.d.periodic[face / 2]) return -1;
  if (type != kBcOutflow && type != kBcReflect && type != kBcFixedFace) return -1;
  m.d.bc[face] = type;
  return 0;
}

void orc_mesh_refine_box(void *h, const double *box, int level) {
  static_cast<OMesh *>(h)->m->AddRefinementBox(box, level);
}

void orc_mesh_refine_sphere(void *h, const double *c, double r, int level) {
  static_cast<OMesh *>(h)->m->AddRefinementSphere(c, r, level);
}

void orc_mesh_finalize(void *h) { static_cast<OMesh *>(h)->m->Finalize(); }

void orc_mesh_free(void *h) { delete static_cast<OMesh *>(h); }

int orc_mesh_nblocks(void *h) { return (int)static_cast<OMesh *>(h)->m->blocks.size(); }

int orc_mesh_multilevel(void *h) { return static_cast<OMesh *>(h)->m->multilevel ? 1 : 0; }

// out: level, l1, l2, l3 ; x: xmin1..3, xmax1..3
void orc_mesh_block(void *h, int gid, std::int64_t *out, double *x) {
  const Block &b = static_cast<OMesh *>(h)->m->blocks.at(gid);
  out[0] = b.loc.level;
  for (int k = 0; k < 3; ++k) {
    out[1 + k] = b.loc.l[k];
    x[k] = b.xmin[k];
    x[3 + k] = b.xmax[k];
  }
}

int orc_mesh_nneighbors(void *h, int gid) {
  return (int)static_cast<OMesh *>(h)->m->blocks.at(gid).neighbors.size();
}

// out: gid, level, origin l1..3, offsets 1..3
void orc_mesh_neighbor(void *h, int gid, int n, std::int64_t *out) {
  const NeighborBlock &nb = static_cast<OMesh *>(h)->m->blocks.at(gid).neighbors.at(n);
  out[0] = nb.gid;
  out[1] = nb.origin_loc.level;
  for (int k = 0; k < 3; ++k) {
    out[2 + k] = nb.origin_loc.l[k];
    out[5 + k] = nb.off[k];
  }
}

// out: cellbounds n[3], ng[3], c_cellbounds n[3], ng[3]
void orc_mesh_shapes(void *h, int *out) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  for (int k = 0; k < 3; ++k) {
    out[k] = m.cellbounds.n[k];
    out[3 + k] = m.cellbounds.ng[k];
    out[6 + k] = m.c_cellbounds.n[k];
    out[9 + k] = m.c_cellbounds.ng[k];
  }
}

// Index boxes of every boundary region of block gid: records of 7 ints {kind, i_lo, i_hi,
// j_lo, j_hi, k_lo, k_hi}; kind 0 send, 1 receive, 2 restriction before send, 3 prolongation
// after set, 4 restriction after set (in the order the reference builds its caches).
int orc_mesh_region_boxes(void *h, int gid, int *out, int max) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  const Block &b = m.blocks.at(gid);
  int n = 0;
  auto put = [&](int kind, const Idx6 &x) {
    if (n >= max) return;
    int *o = out + 7 * n;
    o[0] = kind;
    o[1] = x.s[5];
    o[2] = x.e[5];
    o[3] = x.s[4];
    o[4] = x.e[4];
    o[5] = x.s[3];
    o[6] = x.e[3];
    ++n;
  };
  for (const auto &nb : b.neighbors) {
    put(0, CalcIndices(m, b, nb, 1, nb.IsCell() ? IRT::InteriorSend : IRT::BoundaryInteriorSend,
                       false));
    put(1, CalcIndices(m, b, nb, 1, nb.IsCell() ? IRT::InteriorRecv : IRT::BoundaryExteriorRecv,
                       false));
    if (!m.multilevel) continue;
    const bool coarser = nb.origin_loc.level < b.loc.level;
    if (coarser) {
      put(2, CalcIndices(m, b, nb, 1, IRT::BoundaryInteriorSend, true));
      put(3, CalcIndices(m, b, nb, 1, IRT::BoundaryExteriorRecv, true));
    } else if (b.loc.level > 0 && b.has_coarser_neighbor) {
      put(4, CalcIndices(m, b, nb, 1, IRT::BoundaryExteriorRecv, true));
    }
    // flux correction of a face flux variable (element F(d+1)): 5 restrict+send box in coarse
    // index space, 6 receive box
    const int d = FluxCorrection::FaceDir(nb);
    if (d < 0) continue;
    if (nb.loc.level == b.loc.level - 1)
      put(5, CalcIndices(m, b, nb, 1, IRT::BoundaryInteriorSend, false, d + 1, true));
    if (nb.loc.level - 1 == b.loc.level)
      put(6, CalcIndices(m, b, nb, 1, IRT::BoundaryExteriorRecv, false, d + 1, true));
  }
  return n;
}

// Number of buffer elements actually used by one Send over nfields variables, plus the
// allocated (GetBufferSize) total in *allocated.
std::int64_t orc_exchange_counts(void *h, int nfields, const int *ncomp, std::int64_t *allocated) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  std::int64_t used = 0, alloc = 0;
  for (const auto &b : m.blocks)
    for (int f = 0; f < nfields; ++f)
      for (const auto &nb : b.neighbors) {
        IRT st = nb.IsCell() ? IRT::InteriorSend : IRT::BoundaryInteriorSend;
        used += CalcIndices(m, b, nb, ncomp[f], st, false).size();
        alloc += GetBufferSize(m, nb, ncomp[f]);
      }
  if (allocated) *allocated = alloc;
  return used;
}

// Send + Set (+ prolongation: 0 none, 1 minmod, 2 linear) over nfields variables, repeated nrep
// times. data[f]: [nblocks][ncomp][E]; coarse[f]: [nblocks][ncomp][cE] or null for single level.
void orc_exchange(void *h, int nfields, const int *ncomp, double **data, double **coarse,
                  int prolong, int nrep) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  std::vector<Field> fields(nfields);
  const std::int64_t E = m.cellbounds.size(), cE = m.c_cellbounds.size();
  const std::int64_t nb = m.blocks.size();
  for (int f = 0; f < nfields; ++f) {
    fields[f].ncomp = ncomp[f];
    fields[f].data.assign(data[f], data[f] + nb * ncomp[f] * E);
    if (m.multilevel) fields[f].coarse.assign(coarse[f], coarse[f] + nb * ncomp[f] * cE);
  }
  Exchange ex;
  ex.Build(m, fields);
  for (int r = 0; r < nrep; ++r) {
    ex.Send(fields);
    ex.Set(fields);
    if (prolong) ex.Prolongate(fields, prolong == 1);
  }
  for (int f = 0; f < nfields; ++f) {
    std::memcpy(data[f], fields[f].data.data(), sizeof(double) * nb * ncomp[f] * E);
    if (m.multilevel && coarse[f])
      std::memcpy(coarse[f], fields[f].coarse.data(), sizeof(double) * nb * ncomp[f] * cE);
  }
}

// The boundary exchange with physical boundary conditions (Exchange::WithBoundaryConditions,
// minmod prolongation), repeated nrep times; vector[f] marks Metadata::Vector fields.
void orc_exchange_bc(void *h, int nfields, const int *ncomp, const int *vector, double **data,
                     double **coarse, int nrep) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  std::vector<Field> fields(nfields);
  const std::int64_t E = m.cellbounds.size(), cE = m.c_cellbounds.size();
  const std::int64_t nb = m.blocks.size();
  for (int f = 0; f < nfields; ++f) {
    fields[f].ncomp = ncomp[f];
    fields[f].vector = vector[f] != 0;
    fields[f].data.assign(data[f], data[f] + nb * ncomp[f] * E);
    if (m.multilevel) fields[f].coarse.assign(coarse[f], coarse[f] + nb * ncomp[f] * cE);
  }
  Exchange ex;
  ex.Build(m, fields);
  for (int r = 0; r < nrep; ++r) ex.WithBoundaryConditions(fields);
  for (int f = 0; f < nfields; ++f) {
    std::memcpy(data[f], fields[f].data.data(), sizeof(double) * nb * ncomp[f] * E);
    if (m.multilevel && coarse[f])
      std::memcpy(coarse[f], fields[f].coarse.data(), sizeof(double) * nb * ncomp[f] * cE);
  }
}

// Physical boundary conditions alone on data (coarse = 0) or coarse buffers (coarse = 1).
void orc_apply_bc(void *h, int nfields, const int *ncomp, const int *vector, double **arrays,
                  int coarse) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  std::vector<Field> fields(nfields);
  const std::int64_t E = coarse ? m.c_cellbounds.size() : m.cellbounds.size();
  const std::int64_t nb = m.blocks.size();
  for (int f = 0; f < nfields; ++f) {
    fields[f].ncomp = ncomp[f];
    fields[f].vector = vector[f] != 0;
    std::vector<double> &v = coarse ? fields[f].coarse : fields[f].data;
    v.assign(arrays[f], arrays[f] + nb * ncomp[f] * E);
  }
  ApplyBoundaryConditions(m, fields, coarse != 0);
  for (int f = 0; f < nfields; ++f) {
    const std::vector<double> &v = coarse ? fields[f].coarse : fields[f].data;
    std::memcpy(arrays[f], v.data(), sizeof(double) * nb * ncomp[f] * E);
  }
}

// Restrict block interiors: coarse(c,k,j,i) over the whole coarse interior of every block.
void orc_restrict_interior(void *h, int ncomp, const double *data, double *coarse) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  Field f;
  f.ncomp = ncomp;
  const std::int64_t nb = m.blocks.size();
  f.data.assign(data, data + nb * ncomp * m.cellbounds.size());
  f.coarse.assign(coarse, coarse + nb * ncomp * m.c_cellbounds.size());
  const Shape &cs = m.c_cellbounds;
  for (const auto &b : m.blocks)
    for (int c = 0; c < ncomp; ++c)
      for (int k = cs.s(2); k <= cs.e(2); ++k)
        for (int j = cs.s(1); j <= cs.e(1); ++j)
          for (int i = cs.s(0); i <= cs.e(0); ++i) RestrictCell(m, b, f, c, k, j, i);
  std::memcpy(coarse, f.coarse.data(), sizeof(double) * f.coarse.size());
}

// Prolongate one coarse cell per listed index (gid, c, k, j, i) from coarse into data.
void orc_prolongate_cells(void *h, int ncomp, double *data, const double *coarse, int ncells,
                          const int *cells, int minmod) {
  const Mesh &m = *static_cast<OMesh *>(h)->m;
  Field f;
  f.ncomp = ncomp;
  const std::int64_t nb = m.blocks.size();
  f.data.assign(data, data + nb * ncomp * m.cellbounds.size());
  f.coarse.assign(coarse, coarse + nb * ncomp * m.c_cellbounds.size());
  for (int q = 0; q < ncells; ++q) {
    const int *c = cells + 5 * q;
    ProlongateCell(m, m.blocks.at(c[0]), f, c[1], c[2], c[3], c[4], minmod != 0);
  }
  std::memcpy(data, f.data.data(), sizeof(double) * f.data.size());
}

void orc_weno5z(int n, const double *q, double *ql, double *qr) {
  for (int i = 0; i < n; ++i)
    WENO5Z(q[5 * i], q[5 * i + 1], q[5 * i + 2], q[5 * i + 3], q[5 * i + 4], ql[i], qr[i]);
}

void orc_linear(int n, const double *q, double *ql, double *qr) {
  for (int i = 0; i < n; ++i) Linear(q[3 * i], q[3 * i + 1], q[3 * i + 2], ql[i], qr[i]);
}

// in: 8 per face (uxl uxr uyl uyr uzl uzr upl upr); out: 5 per face (sl sr fux fuy fuz)
void orc_lr_to_flux(int n, const double *in, double *out) {
  for (int i = 0; i < n; ++i) {
    const double *a = in + 8 * i;
    double *o = out + 5 * i;
    lr_to_flux(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], o[0], o[1], o[2], o[3], o[4]);
  }
}

// in: 4 per face (upl upr ql qr); out: the scalar flux with the face's Davis speeds
void orc_scalar_flux(int n, const double *in, double *out) {
  for (int i = 0; i < n; ++i) {
    const double *a = in + 4 * i;
    const double sl = std::min(std::min(a[0], a[1]), 0.0);
    const double sr = std::max(std::max(a[0], a[1]), 0.0);
    out[i] = qflux(a[0], a[1], a[2], a[3], sl, sr);
  }
}

// ---- Burgers benchmark ----
struct OBurgers {
  Burgers b;
};

void *orc_burgers_new(void *mesh, int num_scalars, double cfl, int recon) {
  auto *o = new OBurgers;
  o->b.Init(*static_cast<OMesh *>(mesh)->m, num_scalars, cfl, recon);
  return o;
}

void orc_burgers_free(void *h) { delete static_cast<OBurgers *>(h); }

// Seeded sum-of-sinusoids initial condition: for each component c and mode m = 0..3,
// amplitude A ~ U(0.1, 1), phase phi ~ U(0, 2 pi) drawn in (c, m, A, phi) order from
// SplitMix64(seed); mode m varies along direction (c + m) % 3 with wavenumber m + 1:
//   U_c(x) = sum_m A sin(2 pi (m + 1) x_{(c+m)%3} + phi)   over block interiors.
void orc_burgers_pgen_sinusoid(void *h, std::uint64_t seed) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  SplitMix64 rng{seed};
  std::vector<double> A(B.V * 4), P(B.V * 4);
  for (int c = 0; c < B.V; ++c)
    for (int m = 0; m < 4; ++m) {
      A[4 * c + m] = 0.1 + 0.9 * rng.uniform();
      P[4 * c + m] = 2.0 * M_PI * rng.uniform();
    }
  const Shape &s = B.m->cellbounds;
  for (int b = 0; b < (int)B.m->blocks.size(); ++b) {
    const Coords &co = B.m->blocks[b].coords;
    for (int c = 0; c < B.V; ++c)
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i) {
            const double x[3] = {co.Xc(0, i), co.Xc(1, j), co.Xc(2, k)};
            double v = 0.0;
            for (int m = 0; m < 4; ++m)
              v += A[4 * c + m] * std::sin(2.0 * M_PI * (m + 1) * x[(c + m) % 3] + P[4 * c + m]);
            B.u(B.base.U.data, b, c, k, j, i) = v;
          }
  }
}

// The reference's own tanh/exp problem generator (benchmarks/burgers/parthenon_app_inputs.cpp:34-79)
void orc_burgers_pgen_vibe(void *h) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  const Shape &s = B.m->cellbounds;
  for (int b = 0; b < (int)B.m->blocks.size(); ++b) {
    const Coords &co = B.m->blocks[b].coords;
    for (int k = s.s(2); k <= s.e(2); ++k)
      for (int j = s.s(1); j <= s.e(1); ++j)
        for (int i = s.s(0); i <= s.e(0); ++i) {
          const double x = co.Xc(0, i), y = co.Xc(1, j), z = co.Xc(2, k);
          const double qx = (std::tanh(-20.0 * x) * std::cos(M_PI * x) + 1.0) *
                            std::exp(-30.0 * y * y) * std::exp(-30.0 * z * z);
          const double qy = (std::sin(M_PI * y) + 0.2) * std::exp(-30.0 * x * x) *
                            std::exp(-30.0 * z * z);
          const double qz = (std::tanh(-20. * z) * std::cos(M_PI * z) + 0.5) *
                            std::exp(-30.0 * x * x) * std::exp(-30.0 * y * y);
          B.u(B.base.U.data, b, 0, k, j, i) = qx;
          B.u(B.base.U.data, b, 1, k, j, i) = qy;
          B.u(B.base.U.data, b, 2, k, j, i) = qz;
          for (int n = 3; n < B.V; ++n) {
            double q = 1;
            if (std::abs(x) < 0.025 && std::abs(y) < 0.15 && std::abs(z) < 0.025) q += 10.0;
            B.u(B.base.U.data, b, n, k, j, i) = q;
          }
        }
  }
}

void orc_burgers_initialize(void *h) { static_cast<OBurgers *>(h)->b.Initialize(); }

// EvolutionDriver::Execute's loop (driver.cpp:98-151): at most n cycles while KeepGoing().
int orc_burgers_cycles(void *h, int n) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  int q = 0;
  for (; q < n && B.KeepGoing(); ++q) B.Cycle();
  return q;
}

// <parthenon/time> tlim, dt_init, dt_init_force, dt_user, dt_force, dt_floor, dt_ceil, nlim;
// call before orc_burgers_initialize.
void orc_burgers_set_time_limits(void *h, const double *v, int dt_init_force, int nlim) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  B.lim.tlim = v[0];
  B.lim.dt_init = v[1];
  B.lim.dt_init_force = dt_init_force != 0;
  B.lim.dt_user = v[2];
  B.lim.dt_force = v[3];
  B.lim.dt_floor = v[4];
  B.lim.dt_ceil = v[5];
  B.lim.nlim = nlim;
}
int orc_burgers_ncycle(void *h) { return static_cast<OBurgers *>(h)->b.ncycle; }

// Stage-level entry points for kernel parity tests on the "base" container.
void orc_burgers_fluxes(void *h) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  B.CalculateFluxes(B.base);
}
void orc_burgers_fluxdiv(void *h) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  B.FluxDivergence(B.base, B.dudt);
}
void orc_burgers_derived(void *h) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  B.CalculateDerived(B.base);
}
double orc_burgers_estimate_dt(void *h) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  return B.EstimateTimestep(B.base);
}

double orc_burgers_dt(void *h) { return static_cast<OBurgers *>(h)->b.dt; }
double orc_burgers_time(void *h) { return static_cast<OBurgers *>(h)->b.time; }

static std::vector<double> *Select(Burgers &B, int which) {
  switch (which) {
    case 0: return &B.base.U.data;
    case 1: return &B.one.U.data;
    case 2: return &B.derived;
    case 3: return &B.dudt.U.data;
    case 4: return &B.base.flux[0];
    case 5: return &B.base.flux[1];
    case 6: return &B.base.flux[2];
    default: return nullptr;
  }
}

// which: 0 base U, 1 stage U, 2 derived, 3 dudt, 4..6 base fluxes x/y/z
std::int64_t orc_burgers_size(void *h, int which) {
  auto *v = Select(static_cast<OBurgers *>(h)->b, which);
  return v ? (std::int64_t)v->size() : -1;
}
void orc_burgers_get(void *h, int which, double *out) {
  auto *v = Select(static_cast<OBurgers *>(h)->b, which);
  std::memcpy(out, v->data(), sizeof(double) * v->size());
}
void orc_burgers_set(void *h, int which, const double *in) {
  auto *v = Select(static_cast<OBurgers *>(h)->b, which);
  std::memcpy(v->data(), in, sizeof(double) * v->size());
}

void orc_burgers_set_flux_correction(void *h, int on) {
  static_cast<OBurgers *>(h)->b.flux_correction = on != 0;
}

// Sum over blocks and interior cells of U * cell volume, per component (out[V]), in block /
// k / j / i order: the discrete conserved totals.
void orc_burgers_totals(void *h, double *out) {
  Burgers &B = static_cast<OBurgers *>(h)->b;
  const Shape &s = B.m->cellbounds;
  for (int n = 0; n < B.V; ++n) {
    double acc = 0.0;
    for (int b = 0; b < (int)B.m->blocks.size(); ++b) {
      const double vol = B.m->blocks[b].coords.vol;
      for (int k = s.s(2); k <= s.e(2); ++k)
        for (int j = s.s(1); j <= s.e(1); ++j)
          for (int i = s.s(0); i <= s.e(0); ++i) acc += B.u(B.base.U.data, b, n, k, j, i) * vol;
    }
    out[n] = acc;
  }
}

double orc_burgers_mass_history(void *h, const double *xmn, const double *xmx, double vol) {
  return static_cast<OBurgers *>(h)->b.MassHistory(xmn, xmx, vol);
}

// ---- geometric multigrid Poisson solver (orc_gmg.hpp) ----
struct OGmg {
  orc::gmg::Solver s;
  std::vector<double> u, b;
  explicit OGmg(int N) : s(N), u(s.lv[0].Size(), 0.0), b(s.lv[0].Size(), 0.0) {}
  // interior [k][j][i] of a finest-level array
  void Get(const std::vector<double> &a, double *out) const {
    const auto &L = s.lv[0];
    std::int64_t q = 0;
    for (int k = 0; k < L.n; ++k)
      for (int j = 0; j < L.n; ++j)
        for (int i = 0; i < L.n; ++i) out[q++] = a[L.Idx(k, j, i)];
  }
  void Set(std::vector<double> &a, const double *in) {
    const auto &L = s.lv[0];
    std::int64_t q = 0;
    for (int k = 0; k < L.n; ++k)
      for (int j = 0; j < L.n; ++j)
        for (int i = 0; i < L.n; ++i) a[L.Idx(k, j, i)] = in[q++];
  }
};

void *orc_gmg_new(int N) { return new OGmg(N); }
void orc_gmg_free(void *h) { delete static_cast<OGmg *>(h); }
int orc_gmg_levels(void *h) { return (int)static_cast<OGmg *>(h)->s.lv.size(); }
void orc_gmg_set_params(void *h, double omega, int pre, int post) {
  auto &s = static_cast<OGmg *>(h)->s;
  s.omega = omega;
  s.pre = pre;
  s.post = post;
}
void orc_gmg_rhs_blobs(void *h, std::uint64_t seed) {
  OGmg &g = *static_cast<OGmg *>(h);
  g.s.SetRhsBlobs(seed);
  g.b = g.s.lv[0].b;
}
void orc_gmg_set_rhs(void *h, const double *in) { static_cast<OGmg *>(h)->Set(static_cast<OGmg *>(h)->b, in); }
void orc_gmg_get_rhs(void *h, double *out) { static_cast<OGmg *>(h)->Get(static_cast<OGmg *>(h)->b, out); }
void orc_gmg_set_u(void *h, const double *in) { static_cast<OGmg *>(h)->Set(static_cast<OGmg *>(h)->u, in); }
void orc_gmg_get_u(void *h, double *out) { static_cast<OGmg *>(h)->Get(static_cast<OGmg *>(h)->u, out); }
// A x on the finest level (FixedFace 0 ghosts), interior arrays in and out
void orc_gmg_apply(void *h, const double *x, double *out) {
  OGmg &g = *static_cast<OGmg *>(h);
  std::vector<double> xa(g.s.lv[0].Size(), 0.0), oa(g.s.lv[0].Size(), 0.0);
  g.Set(xa, x);
  g.s.Apply(0, xa, oa);
  g.Get(oa, out);
}
int orc_gmg_solve(void *h, double tol, int maxit) {
  OGmg &g = *static_cast<OGmg *>(h);
  return g.s.Solve(g.u, g.b, tol, maxit);
}
int orc_gmg_history(void *h, double *out, int max) {
  const auto &hist = static_cast<OGmg *>(h)->s.history;
  const int n = (int)hist.size() < max ? (int)hist.size() : max;
  for (int q = 0; q < n; ++q) out[q] = hist[q];
  return (int)hist.size();
}

// transfer operators between finest-level l and l+1 on interior arrays (for their pins)
void orc_gmg_restrict(void *h, int l, const double *fine, double *coarse) {
  OGmg &g = *static_cast<OGmg *>(h);
  auto &F = g.s.lv[l], &C = g.s.lv[l + 1];
  std::int64_t q = 0;
  for (int k = 0; k < F.n; ++k)
    for (int j = 0; j < F.n; ++j)
      for (int i = 0; i < F.n; ++i) F.r[F.Idx(k, j, i)] = fine[q++];
  g.s.Restrict(l);
  q = 0;
  for (int k = 0; k < C.n; ++k)
    for (int j = 0; j < C.n; ++j)
      for (int i = 0; i < C.n; ++i) coarse[q++] = C.b[C.Idx(k, j, i)];
}
void orc_gmg_prolong_add(void *h, int l, const double *coarse, double *fine) {
  OGmg &g = *static_cast<OGmg *>(h);
  auto &F = g.s.lv[l], &C = g.s.lv[l + 1];
  std::int64_t q = 0;
  for (int k = 0; k < C.n; ++k)
    for (int j = 0; j < C.n; ++j)
      for (int i = 0; i < C.n; ++i) C.x[C.Idx(k, j, i)] = coarse[q++];
  q = 0;
  for (int k = 0; k < F.n; ++k)
    for (int j = 0; j < F.n; ++j)
      for (int i = 0; i < F.n; ++i) F.x[F.Idx(k, j, i)] = fine[q++];
  g.s.ProlongAdd(l);
  q = 0;
  for (int k = 0; k < F.n; ++k)
    for (int j = 0; j < F.n; ++j)
      for (int i = 0; i < F.n; ++i) fine[q++] = F.x[F.Idx(k, j, i)];
}

// ---- adaptive remesh ----
// The refinement state of one OMesh. Remeshing replaces the OMesh's mesh, so every orc_mesh_*
// query afterwards sees the new block list.
struct OAmr {
  OMesh *om;
  Adaptive a;
  std::unique_ptr<Mesh> pending;
};

void *orc_amr_new(void *mesh, int numlevel, int deref_threshold) {
  auto *o = new OAmr;
  o->om = static_cast<OMesh *>(mesh);
  o->a.Init(*o->om->m, numlevel, deref_threshold);
  return o;
}
void orc_amr_free(void *h) { delete static_cast<OAmr *>(h); }

void orc_amr_add_criterion(void *h, int method, int comp, double refine_tol,
                           double derefine_tol, int max_level) {
  Criterion c;
  c.method = method;
  c.comp = comp;
  c.refine_tol = refine_tol;
  c.derefine_tol = derefine_tol;
  c.max_level = max_level;
  static_cast<OAmr *>(h)->a.AddCriterion(c);
}

int orc_amr_root_level(void *h) { return static_cast<OAmr *>(h)->a.root_level; }

// largest derivative measure of component comp of every block (out[nblocks])
void orc_amr_derivative_max(void *h, int method, int ncomp, int comp, const double *data,
                            double *out) {
  const Mesh &m = *static_cast<OAmr *>(h)->om->m;
  const std::int64_t sz = m.cellbounds.size();
  for (int b = 0; b < (int)m.blocks.size(); ++b)
    out[b] = DerivativeMax(m, data + ((std::int64_t)b * ncomp + comp) * sz, method);
}

// Refinement::Tag on field data [block][comp][k][j][i]
void orc_amr_tag(void *h, int ncomp, const double *data) {
  OAmr &o = *static_cast<OAmr *>(h);
  Field f;
  f.ncomp = ncomp;
  f.data.assign(data, data + o.om->m->blocks.size() * ncomp * o.om->m->cellbounds.size());
  o.a.Tag(*o.om->m, f);
}

void orc_amr_set_refinement(void *h, int gid, int tag) {
  OAmr &o = *static_cast<OAmr *>(h);
  o.a.SetRefinement(*o.om->m, gid, tag);
}

void orc_amr_flags(void *h, int *flag, int *count) {
  OAmr &o = *static_cast<OAmr *>(h);
  for (size_t b = 0; b < o.a.flag.size(); ++b) {
    flag[b] = o.a.flag[b];
    count[b] = o.a.count[b];
  }
}

// UpdateMeshBlockTree: the block count of the new list, or 0 when the tree did not change.
int orc_amr_update(void *h, int *nnew, int *ndel) {
  OAmr &o = *static_cast<OAmr *>(h);
  std::unique_ptr<Mesh> tree(new Mesh(o.om->m->d));
  o.a.UpdateTree(*o.om->m, *tree, *nnew, *ndel);
  o.pending.reset();
  if (*nnew == 0 && *ndel == 0) return 0;
  o.pending = MeshFromTree(*tree);
  return (int)o.pending->blocks.size();
}

// Moves one field (data of the current mesh) onto the pending mesh's block list.
void orc_amr_transfer(void *h, int ncomp, const double *old_data, double *new_data) {
  OAmr &o = *static_cast<OAmr *>(h);
  Field of, nf;
  of.ncomp = ncomp;
  of.data.assign(old_data, old_data + o.om->m->blocks.size() * ncomp * o.om->m->cellbounds.size());
  TransferField(*o.om->m, of, *o.pending, nf);
  std::memcpy(new_data, nf.data.data(), sizeof(double) * nf.data.size());
}

// Makes the pending mesh current (flags of moved blocks carried, new blocks at 0).
void orc_amr_commit(void *h) {
  OAmr &o = *static_cast<OAmr *>(h);
  CarryFlags(Origins(*o.om->m, *o.pending), o.a);
  o.om->m = std::move(o.pending);
}

// Refine / derefine one location of the current tree directly (Tree::Refine, Tree::Derefine)
// and re-finalize; returns the blocks added or removed.
int orc_mesh_tree_refine(void *mesh, int level, const std::int64_t *l) {
  Mesh &m = *static_cast<OMesh *>(mesh)->m;
  Loc a;
  a.level = level;
  for (int k = 0; k < 3; ++k) a.l[k] = l[k];
  const int n = m.Refine(a);
  m.multilevel = true;
  m.Finalize();
  return n;
}
int orc_mesh_tree_derefine(void *mesh, int level, const std::int64_t *l) {
  Mesh &m = *static_cast<OMesh *>(mesh)->m;
  Loc a;
  a.level = level;
  for (int k = 0; k < 3; ++k) a.l[k] = l[k];
  const int n = m.Derefine(a);
  m.Finalize();
  return n;
}
int orc_mesh_is_leaf(void *mesh, int level, const std::int64_t *l) {
  Mesh &m = *static_cast<OMesh *>(mesh)->m;
  Loc a;
  a.level = level;
  for (int k = 0; k < 3; ++k) a.l[k] = l[k];
  return m.leaves.count(a) ? 1 : 0;
}

// Adaptive Burgers: n cycles of the benchmark with tagging and remeshing after each; returns
// the number of cycles after which the mesh changed. The OBurgers and the OMesh it was built
// on are updated in place.
int orc_amr_burgers_cycles(void *amr, void *burgers, int n) {
  OAmr &o = *static_cast<OAmr *>(amr);
  Burgers &B = static_cast<OBurgers *>(burgers)->b;
  int changed = 0;
  for (int q = 0; q < n; ++q) changed += AdaptiveCycle(o.om->m, B, o.a) ? 1 : 0;
  return changed;
}

// The Poisson problem on the leaf grid of a (possibly refined) mesh: x and out hold the
// interior cells of every block in block order, then [k][j][i]; Df: the face field D as
// [block][3] bricks (component d = D on the lower face along d), or null for the constant D.
void orc_poisson_leaf_apply(void *mesh, double D, double alpha, const double *Df,
                            const double *x, double *out) {
  const Mesh &m = *static_cast<OMesh *>(mesh)->m;
  poisson::LeafProblem P(m, D, alpha);
  P.Df = Df;
  std::vector<double> xv(x, x + P.Cells()), ov;
  P.Apply(xv, ov);
  std::memcpy(out, ov.data(), sizeof(double) * ov.size());
}

// BiCGSTAB on the leaf grid from x (in/out) to RMS residual tol; returns the iterations or -1.
int orc_poisson_leaf_solve(void *mesh, double D, double alpha, const double *Df,
                           const double *b, double *x, double tol, int maxit, double *rms) {
  const Mesh &m = *static_cast<OMesh *>(mesh)->m;
  poisson::LeafProblem P(m, D, alpha);
  P.Df = Df;
  std::vector<double> bv(b, b + P.Cells()), xv(x, x + P.Cells());
  const int it = P.Solve(bv, xv, tol, maxit, rms);
  std::memcpy(x, xv.data(), sizeof(double) * xv.size());
  return it;
}

}  // extern "C"
