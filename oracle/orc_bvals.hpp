// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU restatement of the reference's boundary-buffer path for cell-centred variables:
// BndInfo / ProResInfo index ranges, buffer sizes, pack (SendBoundBufs), unpack
// (SetBounds), cell-average restriction and minmod prolongation (ProlongateBounds).
//
// Restated from (paths relative to the reference root):
//   src/bvals/comms/bnd_info.cpp:105-252   CalcIndices
//   src/bvals/comms/bnd_info.cpp:254-271   GetBufferSize
//   src/bvals/comms/bnd_info.cpp:273-448   BndInfo / ProResInfo constructors
//   src/bvals/comms/bvals_utils.hpp:38-64  Send/Receive buffer keys
//   src/bvals/comms/boundary_communication.cpp:49-159,251-348,361-393  Send/Set/Prolongate
//   src/mesh/forest/block_ownership.cpp:89-141  index-range masks (CC reduces to all-true)
//   src/utils/indexer.hpp:70-190           6-D indexer and IsActive
//   src/prolong_restrict/pr_ops.hpp:78-164 GetGridSpacings, GradMinMod, RestrictAverage
//   src/prolong_restrict/pr_ops.hpp:166-289  ProlongateSharedGeneral (minmod / linear)
//   src/prolong_restrict/pr_loops.hpp:111-159 loop over buffer subsets
//   src/config.hpp.in:86                   SIGN(x) = (x < 0 ? -1 : 1)
#pragma once
#include <cstring>
#include <map>
#include <tuple>

#include "orc_mesh.hpp"

namespace orc {

// A cell-centred variable over all blocks: [block][comp][k][j][i], plus its coarse buffer
// (reference coarse_s) in the c_cellbounds shape when the mesh is multilevel.
struct Field {
  std::string name;
  int ncomp = 1;
  bool fill_ghost = true;
  bool vector = false;  // Metadata::Vector: component v has vector_component v + 1
  std::vector<double> data;
  std::vector<double> coarse;
};

struct Idx6 {  // reference SpatiallyMaskedIndexer6D restricted to what CC fields need
  int s[6], e[6];
  bool mask[27];
  std::int64_t size() const {
    std::int64_t n = 1;
    for (int q = 0; q < 6; ++q) n *= (e[q] - s[q] + 1);
    return n;
  }
  void Decode(std::int64_t idx, int out[6]) const {
    std::int64_t N[6];
    std::int64_t cur = 1;
    for (int q = 5; q >= 0; --q) {
      N[q] = cur;
      cur *= (e[q] - s[q] + 1);
    }
    for (int q = 0; q < 6; ++q) {
      out[q] = (int)(idx / N[q]);
      idx -= out[q] * N[q];
      out[q] += s[q];
    }
  }
  bool IsActive(int k, int j, int i) const {
    int ii = (i == e[5]) - (i == s[5]);
    int jj = (j == e[4]) - (j == s[4]);
    int kk = (k == e[3]) - (k == s[3]);
    return mask[(ii + 1) * 9 + (jj + 1) * 3 + (kk + 1)];
  }
};

enum class IRT { BoundaryInteriorSend, BoundaryExteriorRecv, InteriorSend, InteriorRecv };

// reference: bnd_info.cpp:105-252 for a non-fine variable. el: 0 = cell centre (CC),
// 1..3 = face F1..F3 (the x-, y-, z-face element); flux marks a flux variable, whose regions
// are the shared face itself.
inline Idx6 CalcIndices(const Mesh &m, const Block &b, const NeighborBlock &nb, int ncomp,
                        IRT type, bool prores, int el = 0, bool flux = false) {
  const int g = m.d.nghost;
  Shape shape = (prores || nb.loc.level < b.loc.level) ? m.c_cellbounds : m.cellbounds;
  const int coarse_fac = nb.loc.level > b.loc.level ? 2 : 1;
  int top[3];  // TopologicalOffsetI/J/K (basic_types.hpp:195-203)
  for (int k = 0; k < 3; ++k) top[k] = (el == k + 1) ? 1 : 0;
  int nbn[3];
  for (int k = 0; k < 3; ++k) nbn[k] = m.active[k] ? m.d.bnx[k] / coarse_fac + top[k] : 1;
  int interior_offset = type == IRT::BoundaryInteriorSend ? g : 0;
  int exterior_offset = type == IRT::BoundaryExteriorRecv ? g : 0;
  if (prores) exterior_offset /= 2;
  int s[3], e[3];
  for (int dir = 0; dir < 3; ++dir) {
    // interior bounds of the element (domain.hpp:183-229): faces add one index on the normal
    const int bs = shape.s(dir), be = shape.e(dir) + (m.active[dir] ? top[dir] : 0);
    if (nb.off[dir] == 0) {
      s[dir] = bs;
      e[dir] = be;
      if (b.loc.level < nb.origin_loc.level && m.active[dir]) {
        const int extra = (be - bs + 1) - nbn[dir];
        const std::int64_t par = nb.origin_loc.l[dir] % 2;
        s[dir] += par == 1 ? extra - interior_offset : 0;
        e[dir] -= par == 0 ? extra - interior_offset : 0;
        if (type == IRT::InteriorSend && !prores) {
          s[dir] -= g;
          e[dir] += g;
        }
      }
      if (b.loc.level > nb.origin_loc.level && m.active[dir]) {
        const std::int64_t par = b.loc.l[dir] % 2;
        s[dir] -= par == 1 ? exterior_offset : 0;
        e[dir] += par == 0 ? exterior_offset : 0;
        if (type == IRT::InteriorRecv && !prores) {
          s[dir] -= g;
          e[dir] += g;
        }
      }
      if (prores && m.active[dir] && type == IRT::InteriorRecv) {
        s[dir] -= g / 2;
        e[dir] += g / 2;
      }
    } else if (nb.off[dir] > 0) {
      s[dir] = be + (flux ? 0 : -interior_offset + 1 - top[dir]);
      e[dir] = be + (flux ? 0 : exterior_offset);
    } else {
      s[dir] = bs + (flux ? 0 : -exterior_offset);
      e[dir] = bs + (flux ? 0 : interior_offset - 1 + top[dir]);
    }
  }
  Idx6 out;
  out.s[0] = 0;
  out.e[0] = 0;
  out.s[1] = 0;
  out.e[1] = 0;
  out.s[2] = 0;
  out.e[2] = ncomp - 1;
  out.s[3] = s[2];
  out.e[3] = e[2];
  out.s[4] = s[1];
  out.e[4] = e[1];
  out.s[5] = s[0];
  out.e[5] = e[0];
  // For a cell-centred element GetIndexRangeMaskFromOwnership copies the neighbour's ownership
  // of its own volume (always true) into every entry; for the face element of a face-neighbour
  // flux region the box is one index thick along the normal, so only entries copied from the
  // same volume ownership are read. Either way the mask is all-true.
  for (int q = 0; q < 27; ++q) out.mask[q] = true;
  return out;
}

// RestrictAverage::Do (pr_ops.hpp:105-164) for a face element F(d+1) of a flux variable:
// area-weighted average of the 2x2 fine faces sharing the coarse face, fixed summation order.
inline void RestrictFace(const Mesh &m, const Block &b, const std::vector<double> &fine,
                         std::vector<double> &coarse, int V, int d, int c, int ck, int cj,
                         int ci) {
  const Shape &fs = m.cellbounds, &cs = m.c_cellbounds;
  const bool X[3] = {m.active[0] && d != 0, m.active[1] && d != 1, m.active[2] && d != 2};
  const int i = m.active[0] ? (ci - cs.s(0)) * 2 + fs.s(0) : fs.s(0);
  const int j = m.active[1] ? (cj - cs.s(1)) * 2 + fs.s(1) : fs.s(1);
  const int k = m.active[2] ? (ck - cs.s(2)) * 2 + fs.s(2) : fs.s(2);
  const double area = b.coords.area[d];
  const std::int64_t fslab = (std::int64_t)b.gid * V * fs.size(), cslab = (std::int64_t)b.gid * V * cs.size();
  auto F = [&](int kk, int jj, int ii) {
    return fine[fslab + (((std::int64_t)c * fs.ent(2) + kk) * fs.ent(1) + jj) * fs.ent(0) + ii];
  };
  double vol[2][2][2] = {}, terms[2][2][2] = {};
  for (int ok = 0; ok < 1 + X[2]; ++ok)
    for (int oj = 0; oj < 1 + X[1]; ++oj)
      for (int oi = 0; oi < 1 + X[0]; ++oi) {
        vol[ok][oj][oi] = area;
        terms[ok][oj][oi] = vol[ok][oj][oi] * F(k + ok, j + oj, i + oi);
      }
  const double tvol = ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                      ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
  coarse[cslab + (((std::int64_t)c * cs.ent(2) + ck) * cs.ent(1) + cj) * cs.ent(0) + ci] =
      (((terms[0][0][0] + terms[0][1][0]) + (terms[0][0][1] + terms[0][1][1])) +
       ((terms[1][0][0] + terms[1][1][0]) + (terms[1][0][1] + terms[1][1][1]))) /
      tvol;
}

// Channel key: sender gid, receiver gid, variable index, location index (bvals_utils.hpp:43-64)
using Key = std::tuple<int, int, int, int>;

// Flux correction of a cell variable with fluxes (SendBoundBufs<flxcor_send>, ReceiveBoundBufs /
// SetBounds<flxcor_recv>, boundary_communication.cpp:454-461; ForEachBoundary selection
// loop_utils.hpp:139-162): a block with a coarser face neighbour restricts its face fluxes on
// that face into the flux variable's coarse buffer and sends them; the coarser block overwrites
// its own fluxes on the matching half of the face.
struct FluxCorrection {
  struct Send {
    const Block *b;
    int d;
    Idx6 idx;
    Key key;
  };
  struct Recv {
    int gid, d;
    Idx6 idx;
    Key key;
  };
  const Mesh *m = nullptr;
  std::vector<Send> send;
  std::vector<Recv> recv;
  std::map<Key, std::vector<double>> buffers;

  static int FaceDir(const NeighborBlock &nb) {
    int n = 0, d = -1;
    for (int k = 0; k < 3; ++k)
      if (nb.off[k] != 0) {
        ++n;
        d = k;
      }
    return n == 1 ? d : -1;
  }

  void Build(const Mesh &mesh, int V) {
    m = &mesh;
    send.clear();
    recv.clear();
    buffers.clear();
    if (!mesh.multilevel) return;
    for (const auto &b : mesh.blocks)
      for (const auto &nb : b.neighbors) {
        const int d = FaceDir(nb);
        if (d < 0) continue;
        if (nb.loc.level == b.loc.level - 1) {  // flxcor_send
          Send s;
          s.b = &b;
          s.d = d;
          s.idx = CalcIndices(mesh, b, nb, V, IRT::BoundaryInteriorSend, false, d + 1, true);
          s.key = Key(b.gid, nb.gid, 0, nb.GetIdx());
          buffers[s.key].assign(s.idx.size(), 0.0);
          send.push_back(s);
        }
        if (nb.loc.level - 1 == b.loc.level) {  // flxcor_recv
          Recv r;
          r.gid = b.gid;
          r.d = d;
          r.idx = CalcIndices(mesh, b, nb, V, IRT::BoundaryExteriorRecv, false, d + 1, true);
          r.key = Key(nb.gid, b.gid, 0, nb.GetReverseIdx());
          recv.push_back(r);
        }
      }
  }

  // flux[d]: fine fluxes [block][V][brick]; cflux[d]: coarse buffers [block][V][coarse brick]
  void Run(std::vector<double> *flux, std::vector<double> *cflux, int V) {
    const Shape &fs = m->cellbounds, &cs = m->c_cellbounds;
    for (const auto &s : send) {  // restrict (ProResInfo::GetSend) then pack from coarse_s
      std::vector<double> &buf = buffers.at(s.key);
      for (std::int64_t n = 0; n < s.idx.size(); ++n) {
        int t[6];
        s.idx.Decode(n, t);
        RestrictFace(*m, *s.b, flux[s.d], cflux[s.d], V, s.d, t[2], t[3], t[4], t[5]);
        buf[n] = cflux[s.d][(std::int64_t)s.b->gid * V * cs.size() +
                            (((std::int64_t)t[2] * cs.ent(2) + t[3]) * cs.ent(1) + t[4]) *
                                cs.ent(0) + t[5]];
      }
    }
    for (const auto &r : recv) {  // SetBounds<flxcor_recv>: overwrite the coarse block's fluxes
      const std::vector<double> &buf = buffers.at(r.key);
      for (std::int64_t n = 0; n < r.idx.size(); ++n) {
        int t[6];
        r.idx.Decode(n, t);
        flux[r.d][(std::int64_t)r.gid * V * fs.size() +
                  (((std::int64_t)t[2] * fs.ent(2) + t[3]) * fs.ent(1) + t[4]) * fs.ent(0) +
                  t[5]] = buf[n];
      }
    }
  }
};

// reference: bnd_info.cpp:254-271
inline std::int64_t GetBufferSize(const Mesh &m, const NeighborBlock &nb, int ncomp) {
  const Shape &cb = m.cellbounds;
  const int g = m.d.nghost;
  std::int64_t sz = 1;
  for (int k = 0; k < 3; ++k) sz *= (nb.off[k] == 0 ? cb.ent(k) + 1 : g + 1);
  return sz * ncomp;
}

struct BndInfo {
  Idx6 idx;
  double *var = nullptr;  // block slab of data or coarse
  Shape shape;
  bool same_to_same = false;
  Key key;
};

struct ProRes {
  bool restrict_op = false, prolong_op = false;
  Idx6 idx;
  const Block *block = nullptr;
  int field = 0;
};

inline double *BlockSlab(Field &f, const Shape &sh, int gid, bool coarse) {
  std::vector<double> &v = coarse ? f.coarse : f.data;
  return v.data() + (std::int64_t)gid * f.ncomp * sh.size();
}

inline double &At(double *slab, const Shape &sh, int c, int k, int j, int i) {
  return slab[(((std::int64_t)c * sh.ent(2) + k) * sh.ent(1) + j) * sh.ent(0) + i];
}

inline int SIGN(double x) { return x < 0 ? -1 : 1; }

// reference: pr_ops.hpp:105-164 (3-D, cell-centred, uniform volumes)
inline void RestrictCell(const Mesh &m, const Block &b, Field &f, int c, int ck, int cj, int ci) {
  const Shape &fs = m.cellbounds, &cs = m.c_cellbounds;
  double *fine = BlockSlab(f, fs, b.gid, false);
  double *coarse = BlockSlab(f, cs, b.gid, true);
  const bool X1 = m.active[0], X2 = m.active[1], X3 = m.active[2];
  const int i = X1 ? (ci - cs.s(0)) * 2 + fs.s(0) : fs.s(0);
  const int j = X2 ? (cj - cs.s(1)) * 2 + fs.s(1) : fs.s(1);
  const int k = X3 ? (ck - cs.s(2)) * 2 + fs.s(2) : fs.s(2);
  double vol[2][2][2] = {}, terms[2][2][2] = {};
  for (int ok = 0; ok < 1 + X3; ++ok)
    for (int oj = 0; oj < 1 + X2; ++oj)
      for (int oi = 0; oi < 1 + X1; ++oi) {
        vol[ok][oj][oi] = b.coords.vol;
        terms[ok][oj][oi] = vol[ok][oj][oi] * At(fine, fs, c, k + ok, j + oj, i + oi);
      }
  const double tvol = ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                      ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
  At(coarse, cs, c, ck, cj, ci) =
      (((terms[0][0][0] + terms[0][1][0]) + (terms[0][0][1] + terms[0][1][1])) +
       ((terms[1][0][0] + terms[1][1][0]) + (terms[1][0][1] + terms[1][1][1]))) /
      tvol;
}

// reference: pr_ops.hpp:90-101
inline double GradMinMod(double fc, double fm, double fp, double dxm, double dxp, double &gxm,
                         double &gxp) {
  gxm = (fc - fm) / dxm;
  gxp = (fp - fc) / dxp;
  return 0.5 * (SIGN(gxm) + SIGN(gxp)) * std::min(std::abs(gxm), std::abs(gxp));
}

// reference: pr_ops.hpp:166-283 (cell-centred); minmod=true is ProlongateSharedMinMod,
// minmod=false is ProlongateSharedLinear.
inline void ProlongateCell(const Mesh &m, const Block &b, Field &f, int c, int k, int j, int i,
                           bool minmod) {
  const Shape &fs = m.cellbounds, &cs = m.c_cellbounds;
  double *fine = BlockSlab(f, fs, b.gid, false);
  double *coarse = BlockSlab(f, cs, b.gid, true);
  const bool X1 = m.active[0], X2 = m.active[1], X3 = m.active[2];
  const int fi = X1 ? (i - cs.s(0)) * 2 + fs.s(0) : fs.s(0);
  const int fj = X2 ? (j - cs.s(1)) * 2 + fs.s(1) : fs.s(1);
  const int fk = X3 ? (k - cs.s(2)) * 2 + fs.s(2) : fs.s(2);
  const Coords &cc = b.coarse_coords, &fc_ = b.coords;
  const double fc = At(coarse, cs, c, k, j, i);
  double dxfm[3] = {0, 0, 0}, dxfp[3] = {0, 0, 0}, gm[3] = {0, 0, 0}, gp[3] = {0, 0, 0};
  const int ci[3] = {i, j, k}, fidx[3] = {fi, fj, fk};
  for (int d = 0; d < 3; ++d) {
    if (!m.active[d]) continue;
    const double xm = cc.Xc(d, ci[d] - 1), xc = cc.Xc(d, ci[d]), xp = cc.Xc(d, ci[d] + 1);
    const double dxm = xc - xm, dxp = xp - xc;
    const double fxm = fc_.Xc(d, fidx[d]), fxp = fc_.Xc(d, fidx[d] + 1);
    dxfm[d] = xc - fxm;
    dxfp[d] = fxp - xc;
    int km = k, jm = j, im = i, kp = k, jp = j, ip = i;
    if (d == 0) { im = i - 1; ip = i + 1; }
    if (d == 1) { jm = j - 1; jp = j + 1; }
    if (d == 2) { km = k - 1; kp = k + 1; }
    double gxm, gxp;
    const double gc = GradMinMod(fc, At(coarse, cs, c, km, jm, im), At(coarse, cs, c, kp, jp, ip),
                                 dxm, dxp, gxm, gxp);
    if (minmod) {
      gm[d] = gc;
      gp[d] = gc;
    } else {
      gm[d] = gxm;
      gp[d] = gxp;
    }
  }
  const double g1m = gm[0], g1p = gp[0], g2m = gm[1], g2p = gp[1], g3m = gm[2], g3p = gp[2];
  const double d1m = dxfm[0], d1p = dxfp[0], d2m = dxfm[1], d2p = dxfp[1], d3m = dxfm[2],
               d3p = dxfp[2];
  At(fine, fs, c, fk, fj, fi) = fc - (g1m * d1m + g2m * d2m + g3m * d3m);
  if (X1) At(fine, fs, c, fk, fj, fi + 1) = fc + (g1p * d1p - g2m * d2m - g3m * d3m);
  if (X2) At(fine, fs, c, fk, fj + 1, fi) = fc - (g1m * d1m - g2p * d2p + g3m * d3m);
  if (X2 && X1) At(fine, fs, c, fk, fj + 1, fi + 1) = fc + (g1p * d1p + g2p * d2p - g3m * d3m);
  if (X3) At(fine, fs, c, fk + 1, fj, fi) = fc - (g1m * d1m + g2m * d2m - g3p * d3p);
  if (X3 && X1) At(fine, fs, c, fk + 1, fj, fi + 1) = fc + (g1p * d1p - g2m * d2m + g3p * d3p);
  if (X3 && X2) At(fine, fs, c, fk + 1, fj + 1, fi) = fc - (g1m * d1m - g2p * d2p - g3p * d3p);
  if (X3 && X2 && X1)
    At(fine, fs, c, fk + 1, fj + 1, fi + 1) = fc + (g1p * d1p + g2p * d2p + g3p * d3p);
}

inline void RunRestrict(const Mesh &m, std::vector<Field> &fields, const std::vector<ProRes> &pr) {
  for (const auto &p : pr) {
    if (!p.restrict_op) continue;
    for (std::int64_t n = 0; n < p.idx.size(); ++n) {
      int t[6];
      p.idx.Decode(n, t);
      if (p.idx.IsActive(t[3], t[4], t[5]))
        RestrictCell(m, *p.block, fields[p.field], t[2], t[3], t[4], t[5]);
    }
  }
}

inline void RunProlongate(const Mesh &m, std::vector<Field> &fields, const std::vector<ProRes> &pr,
                          bool minmod) {
  for (const auto &p : pr) {
    if (!p.prolong_op) continue;
    for (std::int64_t n = 0; n < p.idx.size(); ++n) {
      int t[6];
      p.idx.Decode(n, t);
      if (p.idx.IsActive(t[3], t[4], t[5]))
        ProlongateCell(m, *p.block, fields[p.field], t[2], t[3], t[4], t[5], minmod);
    }
  }
}

// ---- physical boundary conditions ----
// reference DoPhysicalBoundary_ (boundary_conditions.cpp:197-211) on the block's boundary flag:
// a face is physical when it lies on the domain boundary of an active, non-periodic dimension.
inline bool PhysicalFace(const Mesh &m, const Block &b, int face) {
  const int d = face / 2;
  if (!m.active[d] || m.d.bc[face] == kBcPeriodic) return false;
  const std::int64_t N = m.NAt(d, b.loc.level);
  return face % 2 == 0 ? b.loc.l[d] == 0 : b.loc.l[d] == N - 1;
}

// reference GenericBC<DIR, SIDE, Outflow|Reflect|FixedFace> (boundary_conditions_generic.hpp:
// 174-255) for one cell-centred FillGhost field of one block: the loop of par_for_bndry over
// IndexDomain::inner_xD / outer_xD (ghost layers along D, the entire range along the other
// dimensions; domain.hpp:184-252), Outflow copying the first / last interior cell (ref),
// Reflect copying the mirror cell (offset - index) and negating the vector component D,
// FixedFace with value 0 setting 2 * 0 - the mirror cell (poisson_package.cpp:60-78).
inline void GenericBC(const Mesh &m, const Block &b, Field &f, bool coarse, int dir, bool inner,
                      int type) {
  const Shape &sh = coarse ? m.c_cellbounds : m.cellbounds;
  double *slab = BlockSlab(f, sh, b.gid, coarse);
  const int ref = inner ? sh.s(dir) : sh.e(dir);
  const int offset = 2 * ref + (inner ? -1 : 1);
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = 0;
    hi[d] = sh.ent(d) - 1;
  }
  if (inner)
    hi[dir] = sh.s(dir) - 1;
  else
    lo[dir] = sh.e(dir) + 1;
  for (int l = 0; l < f.ncomp; ++l) {
    // vector_component (v + 1 for component v of a Vector field, sparse_pack_base.cpp:250-251)
    // equals the direction X1DIR = 1, X2DIR = 2, X3DIR = 3
    const bool reflect = f.vector && l + 1 == dir + 1;
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) {
          int src[3] = {i, j, k};
          const bool mirror = type == kBcReflect || type == kBcFixedFace;
          src[dir] = mirror ? offset - src[dir] : ref;
          const double q = At(slab, sh, l, src[2], src[1], src[0]);
          const double val = 0.0;
          At(slab, sh, l, k, j, i) = type == kBcFixedFace ? 2.0 * val - q
                                     : type == kBcReflect ? (reflect ? -1.0 : 1.0) * q
                                                          : q;
        }
  }
}

// reference ApplyBoundaryConditionsOnCoarseOrFine (boundary_conditions.cpp:36-58), over the
// blocks of the MeshData (ApplyBoundaryConditionsOnCoarseOrFineMD, :82-87): faces in
// BoundaryFace order inner_x1, outer_x1, inner_x2, ..., outer_x3.
inline void ApplyBoundaryConditions(const Mesh &m, std::vector<Field> &fields, bool coarse) {
  for (const auto &b : m.blocks)
    for (int face = 0; face < 6; ++face)
      if (PhysicalFace(m, b, face))
        for (auto &f : fields)
          if (f.fill_ghost) GenericBC(m, b, f, coarse, face / 2, face % 2 == 0, m.d.bc[face]);
}

inline bool HasPhysicalBoundary(const Mesh &m) {
  for (int face = 0; face < 6; ++face)
    if (m.active[face / 2] && m.d.bc[face] != kBcPeriodic) return true;
  return false;
}

// Boundary exchange state for one set of fields on one mesh (reference MeshData caches +
// Mesh::boundary_comm_map). Buffers live here, keyed like the reference channels.
struct Exchange {
  const Mesh *m;
  std::map<Key, std::vector<double>> buffers;
  std::vector<BndInfo> send, recv;
  std::vector<ProRes> send_pr, set_pr;

  void Build(const Mesh &mesh, std::vector<Field> &fields) {
    m = &mesh;
    send.clear();
    recv.clear();
    send_pr.clear();
    set_pr.clear();
    buffers.clear();
    for (const auto &b : mesh.blocks) {
      for (int fi = 0; fi < (int)fields.size(); ++fi) {
        Field &f = fields[fi];
        if (!f.fill_ghost) continue;
        for (const auto &nb : b.neighbors) {
          // send side (GetSendBndInfo + ProResInfo::GetSend)
          IRT st = nb.IsCell() ? IRT::InteriorSend : IRT::BoundaryInteriorSend;
          BndInfo s;
          bool coarser = nb.origin_loc.level < b.loc.level;
          s.shape = coarser ? mesh.c_cellbounds : mesh.cellbounds;
          s.var = BlockSlab(f, s.shape, b.gid, coarser);
          s.idx = CalcIndices(mesh, b, nb, f.ncomp, st, false);
          s.same_to_same = (b.gid == nb.gid) && nb.IsCell();
          s.key = Key(b.gid, nb.gid, fi, nb.GetIdx());
          buffers[s.key].assign(GetBufferSize(mesh, nb, f.ncomp), 0.0);
          send.push_back(s);
          ProRes sp;
          sp.block = &b;
          sp.field = fi;
          if (coarser) {
            sp.restrict_op = true;
            sp.idx = CalcIndices(mesh, b, nb, f.ncomp, IRT::BoundaryInteriorSend, true);
          }
          send_pr.push_back(sp);
          // receive side (GetSetBndInfo + ProResInfo::GetSet)
          IRT rt = nb.IsCell() ? IRT::InteriorRecv : IRT::BoundaryExteriorRecv;
          BndInfo r;
          r.shape = coarser ? mesh.c_cellbounds : mesh.cellbounds;
          r.var = BlockSlab(f, r.shape, b.gid, coarser);
          r.idx = CalcIndices(mesh, b, nb, f.ncomp, rt, false);
          r.same_to_same = (b.gid == nb.gid) && nb.IsCell();
          r.key = Key(nb.gid, b.gid, fi, nb.GetReverseIdx());
          recv.push_back(r);
          ProRes rp;
          rp.block = &b;
          rp.field = fi;
          const bool restricted = b.loc.level > 0 && b.has_coarser_neighbor;
          if (coarser) {
            rp.prolong_op = true;
            rp.idx = CalcIndices(mesh, b, nb, f.ncomp, IRT::BoundaryExteriorRecv, true);
          } else if (restricted) {
            rp.restrict_op = true;
            rp.idx = CalcIndices(mesh, b, nb, f.ncomp, IRT::BoundaryExteriorRecv, true);
          }
          set_pr.push_back(rp);
        }
      }
    }
  }

  // SendBoundBufs: restrict into coarse buffers, then pack (boundary_communication.cpp:49-159)
  void Send(std::vector<Field> &fields) {
    if (m->multilevel) RunRestrict(*m, fields, send_pr);
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t q = 0; q < send.size(); ++q) {
      const BndInfo &s = send[q];
      if (s.same_to_same) continue;
      std::vector<double> &buf = buffers.at(s.key);
      for (std::int64_t n = 0; n < s.idx.size(); ++n) {
        int t[6];
        s.idx.Decode(n, t);
        buf[n] = At(s.var, s.shape, t[2], t[3], t[4], t[5]);
      }
    }
  }

  // SetBounds: unpack, then restrict ghost regions needed by prolongation (:251-348)
  void Set(std::vector<Field> &fields) {
    // regions of one Set write disjoint cells (or the same value from the same source)
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t q = 0; q < recv.size(); ++q) {
      const BndInfo &r = recv[q];
      if (r.same_to_same) continue;
      const std::vector<double> &buf = buffers.at(r.key);
      for (std::int64_t n = 0; n < r.idx.size(); ++n) {
        int t[6];
        r.idx.Decode(n, t);
        if (r.idx.IsActive(t[3], t[4], t[5])) At(r.var, r.shape, t[2], t[3], t[4], t[5]) = buf[n];
      }
    }
    if (m->multilevel) RunRestrict(*m, fields, set_pr);
  }

  // ProlongateBounds (:361-393); ProlongateInternal is the identity for cell-centred fields.
  void Prolongate(std::vector<Field> &fields, bool minmod = true) {
    if (m->multilevel) RunProlongate(*m, fields, set_pr, minmod);
  }

  void All(std::vector<Field> &fields) {
    Send(fields);
    Set(fields);
    Prolongate(fields);
  }

  // AddBoundaryExchangeTasks (boundary_communication.cpp:432-445) and the boundary part of
  // Mesh::CommunicateBoundaries (mesh.cpp:697-705): send, set (+ restriction), coarse-buffer
  // boundary conditions, prolongation, boundary conditions.
  void WithBoundaryConditions(std::vector<Field> &fields, bool minmod = true) {
    Send(fields);
    Set(fields);
    const bool bc = HasPhysicalBoundary(*m);
    if (m->multilevel) {
      if (bc) ApplyBoundaryConditions(*m, fields, true);
      RunProlongate(*m, fields, set_pr, minmod);
    }
    if (bc) ApplyBoundaryConditions(*m, fields, false);
  }
};

}  // namespace orc
