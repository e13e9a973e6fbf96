// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU restatement of the reference's mesh topology for the per-stage meshblock data path:
// logical locations, static refinement with proper nesting, neighbour finding, block
// coordinates and index shapes. Only tests/, the smoke check and the CPU-baseline leg of
// bench.py may load this code; the product library never links it.
//
// Restated from (paths relative to the reference root):
//   src/mesh/forest/tree.cpp:72-134        AddMeshBlock / Refine with proper nesting
//   src/mesh/forest/tree.cpp:139-233       FindNeighbors (same, finer, coarser neighbours)
//   src/mesh/forest/logical_location.cpp:61-79,100-129   symmetrized coords, offsets, IsNeighbor
//   src/mesh/mesh.cpp:1054-1116            static refinement index ranges
//   src/mesh/mesh-gmg.cpp:48-94            NeighborBlock construction (offsets)
//   src/mesh/meshblock.cpp:201-218         cellbounds / c_cellbounds
//   src/coordinates/uniform_cartesian.hpp:30-58   fine and coarse coordinates
//   src/defs.hpp:98-101                    SymmetrizedLogicalToActualPosition
//
// Simplification (stated): the root grid is treated as one Cartesian lattice of base
// blocks with identity logical-coordinate transforms. This is exactly the reference's
// single-tree forest (every configuration used here has a power-of-two block count per
// dimension, equal in all dimensions); multi-tree forests with rotated trees are out of scope.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

struct Loc {
  int level = 0;
  std::int64_t l[3] = {0, 0, 0};
  bool operator<(const Loc &o) const {
    if (level != o.level) return level < o.level;
    for (int d = 2; d >= 0; --d)
      if (l[d] != o.l[d]) return l[d] < o.l[d];
    return false;
  }
  bool operator==(const Loc &o) const {
    return level == o.level && l[0] == o.l[0] && l[1] == o.l[1] && l[2] == o.l[2];
  }
};

inline std::int64_t floordiv2(std::int64_t a) { return a >= 0 ? a / 2 : -((-a + 1) / 2); }

inline Loc Parent(const Loc &a) {
  Loc p;
  p.level = a.level - 1;
  for (int d = 0; d < 3; ++d) p.l[d] = floordiv2(a.l[d]);
  return p;
}

// reference: logical_location.cpp:61-74
inline double IndexToSymmetrizedCoordinate(std::int64_t index, int bloc, std::int64_t nrange) {
  std::int64_t noffset = index - nrange / 2;
  std::int64_t noffset_ceil = index - (nrange + 1) / 2;
  return static_cast<double>(noffset + noffset_ceil + static_cast<std::int64_t>(bloc)) /
         (2.0 * nrange);
}

// One active-dimension index range of a block array (reference IndexShape).
struct Shape {
  int n[3];   // interior cells
  int ng[3];  // ghosts per side
  int ent(int d) const { return n[d] + 2 * ng[d]; }
  int s(int d) const { return ng[d]; }
  int e(int d) const { return ng[d] + n[d] - 1; }
  std::int64_t size() const { return (std::int64_t)ent(0) * ent(1) * ent(2); }
};

struct Coords {  // reference UniformCartesian (uniform_cartesian.hpp:30-58)
  double dx[3], xmin[3], area[3], vol;
  int istart[3];
  double Xc(int d, int idx) const { return xmin[d] + (idx + 0.5) * dx[d]; }
};

struct NeighborBlock {
  int gid;
  Loc loc;         // location in the (wrapped) domain
  Loc origin_loc;  // location as seen from the origin block (unwrapped across periodic faces)
  int off[3];
  bool IsCell() const { return off[0] == 0 && off[1] == 0 && off[2] == 0; }
  int GetIdx() const { return (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1); }
  int GetReverseIdx() const { return (1 - off[0]) + 3 * (1 - off[1]) + 9 * (1 - off[2]); }
};

struct Block {
  int gid;
  Loc loc;
  double xmin[3], xmax[3];
  Coords coords, coarse_coords;
  std::vector<NeighborBlock> neighbors;
  bool has_coarser_neighbor = false;
};

// Boundary condition of a domain face (reference <parthenon/mesh> ix1_bc ... ox3_bc,
// mesh.cpp:1180-1185, default "outflow"; "periodic" on both faces of a periodic dimension).
enum BcType { kBcPeriodic = 0, kBcOutflow = 1, kBcReflect = 2, kBcFixedFace = 3 };

struct MeshDesc {
  int nx[3];
  int bnx[3];
  int nghost;
  double xmin[3], xmax[3];
  int periodic[3];
  int bc[6] = {kBcPeriodic, kBcPeriodic, kBcPeriodic, kBcPeriodic, kBcPeriodic, kBcPeriodic};
};

struct Mesh {
  MeshDesc d;
  int ndim = 3;
  bool active[3];
  std::int64_t nbase[3];
  bool multilevel = false;
  std::set<Loc> leaves, internal;
  std::vector<Block> blocks;
  std::map<Loc, int> gid_of;
  Shape cellbounds, c_cellbounds;

  explicit Mesh(const MeshDesc &desc) : d(desc) {
    ndim = 0;
    for (int k = 0; k < 3; ++k) {
      active[k] = d.nx[k] > 1;
      if (active[k]) ndim = k + 1;
      if (d.nx[k] % d.bnx[k] != 0) throw std::runtime_error("block size must divide mesh");
      nbase[k] = d.nx[k] / d.bnx[k];
    }
    for (std::int64_t k = 0; k < nbase[2]; ++k)
      for (std::int64_t j = 0; j < nbase[1]; ++j)
        for (std::int64_t i = 0; i < nbase[0]; ++i) {
          Loc a;
          a.level = 0;
          a.l[0] = i;
          a.l[1] = j;
          a.l[2] = k;
          leaves.insert(a);
        }
  }

  std::int64_t NAt(int dir, int level) const { return nbase[dir] << level; }

  // Wrap a location into the domain; returns false if it leaves a non-periodic domain.
  bool Wrap(Loc &a) const {
    for (int k = 0; k < 3; ++k) {
      std::int64_t N = NAt(k, a.level);
      if (a.l[k] < 0 || a.l[k] >= N) {
        if (!active[k] || !d.periodic[k]) return false;
        a.l[k] = ((a.l[k] % N) + N) % N;
      }
    }
    return true;
  }

  std::vector<Loc> Daughters(const Loc &a) const {
    std::vector<Loc> out;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < (ndim > 1 ? 2 : 1); ++j)
        for (int k = 0; k < (ndim > 2 ? 2 : 1); ++k) {
          Loc c;
          c.level = a.level + 1;
          c.l[0] = 2 * a.l[0] + i;
          c.l[1] = active[1] ? 2 * a.l[1] + j : 0;
          c.l[2] = active[2] ? 2 * a.l[2] + k : 0;
          out.push_back(c);
        }
    return out;
  }

  // reference: tree.cpp:96-134
  int Refine(const Loc &ref) {
    if (!leaves.count(ref)) return 0;
    leaves.erase(ref);
    internal.insert(ref);
    for (auto &c : Daughters(ref)) leaves.insert(c);
    int added = (1 << ndim) - 1;
    if (ref.level > 0) {
      Loc parent = Parent(ref);
      std::int64_t o[3];
      for (int k = 0; k < 3; ++k) o[k] = ref.l[k] - 2 * parent.l[k];
      for (int kk = 0; kk < (ndim > 2 ? 2 : 1); ++kk)
        for (int jj = 0; jj < (ndim > 1 ? 2 : 1); ++jj)
          for (int ii = 0; ii < 2; ++ii) {
            Loc n = parent;
            n.l[0] += ii + o[0] - 1;
            if (active[1]) n.l[1] += jj + o[1] - 1;
            if (active[2]) n.l[2] += kk + o[2] - 1;
            if (!Wrap(n)) continue;
            added += Refine(n);
          }
    }
    return added;
  }

  // reference: tree.cpp:228-275 (Tree::Derefine with proper nesting). The daughters of
  // `parent` must all be leaves, and no same-level neighbour of a daughter (the 3^ndim cube
  // around it, periodic images included) may be an internal node.
  int Derefine(const Loc &parent) {
    const std::vector<Loc> ds = Daughters(parent);
    for (const Loc &dl : ds) {
      if (!leaves.count(dl)) return 0;
      for (int k = (ndim > 2 ? -1 : 0); k <= (ndim > 2 ? 1 : 0); ++k)
        for (int j = (ndim > 1 ? -1 : 0); j <= (ndim > 1 ? 1 : 0); ++j)
          for (int i = -1; i <= 1; ++i) {
            Loc n = dl;
            n.l[0] += i;
            n.l[1] += j;
            n.l[2] += k;
            if (!Wrap(n)) continue;
            if (internal.count(n)) return 0;
          }
    }
    for (const Loc &dl : ds) leaves.erase(dl);
    internal.erase(parent);
    leaves.insert(parent);
    return static_cast<int>(ds.size()) - 1;
  }

  // reference: forest.cpp:64-105 (Forest::HyperRectangular). The base lattice is a grid of
  // trees of 2^root_level blocks per side; root_level = log2 of the largest power of two
  // dividing the base block count of every dimension (0 when a dimension is inactive).
  // Logical levels of the reference are root_level + the levels used here.
  int RootLevel() const {
    std::int64_t div = std::numeric_limits<int>::max();
    for (int k = 0; k < 3; ++k) {
      if (!active[k]) {
        div = 1;
        continue;
      }
      std::int64_t p = 1;
      while (nbase[k] % (2 * p) == 0) p *= 2;
      div = std::min(div, p);
    }
    int lev = 0;
    while ((std::int64_t(1) << (lev + 1)) <= div) ++lev;
    return lev;
  }

  // reference: tree.cpp:72-94
  void AddMeshBlock(const Loc &a) {
    if (internal.count(a) || leaves.count(a)) return;
    std::vector<Loc> stack;
    Loc p = Parent(a);
    for (int lev = a.level - 1; lev >= 0; --lev) {
      stack.push_back(p);
      if (leaves.count(p)) break;
      p = Parent(p);
    }
    while (!stack.empty()) {
      Refine(stack.back());
      stack.pop_back();
    }
    multilevel = true;
  }

  double ActualPosition(int dir, double u) const {  // defs.hpp:98-101
    return 0.5 * (d.xmin[dir] + d.xmax[dir]) + (u * d.xmax[dir] - u * d.xmin[dir]);
  }

  // reference: mesh.cpp:1063-1081 (GetStaticRefLLIndexRange)
  std::pair<std::int64_t, std::int64_t> StaticRange(int dir, int level, double rmin,
                                                    double rmax) const {
    std::int64_t lxtot = NAt(dir, level);
    std::int64_t lo, hi;
    for (lo = 0; lo < lxtot; lo++) {
      double r = IndexToSymmetrizedCoordinate(lo + 1, 0, lxtot);
      if (ActualPosition(dir, r) > rmin) break;
    }
    for (hi = lo; hi < lxtot; hi++) {
      double r = IndexToSymmetrizedCoordinate(hi + 1, 0, lxtot);
      if (ActualPosition(dir, r) >= rmax) break;
    }
    if (lo % 2 == 1) lo--;
    if (hi % 2 == 0) hi++;
    return {lo, hi};
  }

  // reference: mesh.cpp:1085-1116 (one <parthenon/static_refinementN> block)
  void AddRefinementBox(const double box[6], int level) {
    std::int64_t lmin[3] = {0, 0, 0}, lmax[3] = {1, 1, 1};
    for (int k = 0; k < 3; ++k)
      if (active[k]) {
        auto r = StaticRange(k, level, box[2 * k], box[2 * k + 1]);
        lmin[k] = r.first;
        lmax[k] = r.second;
      }
    for (std::int64_t k = lmin[2]; k < lmax[2]; k += 2)
      for (std::int64_t j = lmin[1]; j < lmax[1]; j += 2)
        for (std::int64_t i = lmin[0]; i < lmax[0]; i += 2) {
          Loc a;
          a.level = level;
          a.l[0] = i;
          a.l[1] = j;
          a.l[2] = k;
          AddMeshBlock(a);
        }
  }

  void BlockExtent(const Loc &a, double *xmn, double *xmx) const {
    for (int k = 0; k < 3; ++k) {
      if (!active[k]) {
        xmn[k] = d.xmin[k];
        xmx[k] = d.xmax[k];
        continue;
      }
      std::int64_t N = NAt(k, a.level);
      xmn[k] = ActualPosition(k, IndexToSymmetrizedCoordinate(a.l[k], 0, N));
      xmx[k] = ActualPosition(k, IndexToSymmetrizedCoordinate(a.l[k], 2, N));
    }
  }

  // Sphere refinement expressed as one reference refinement box per intersecting block.
  void AddRefinementSphere(const double c[3], double r, int level) {
    std::int64_t N[3] = {NAt(0, level), active[1] ? NAt(1, level) : 1,
                         active[2] ? NAt(2, level) : 1};
    for (std::int64_t k = 0; k < N[2]; ++k)
      for (std::int64_t j = 0; j < N[1]; ++j)
        for (std::int64_t i = 0; i < N[0]; ++i) {
          Loc a;
          a.level = level;
          a.l[0] = i;
          a.l[1] = j;
          a.l[2] = k;
          double mn[3], mx[3];
          BlockExtent(a, mn, mx);
          double dist2 = 0;
          for (int q = 0; q < 3; ++q) {
            if (!active[q]) continue;
            double p = std::min(std::max(c[q], mn[q]), mx[q]);
            dist2 += (p - c[q]) * (p - c[q]);
          }
          if (dist2 < r * r) AddMeshBlock(a);
        }
  }

  static std::uint64_t Morton(const Loc &a, int maxlev) {
    std::uint64_t key = 0;
    std::uint64_t x[3];
    for (int k = 0; k < 3; ++k) x[k] = (std::uint64_t)a.l[k] << (maxlev - a.level);
    for (int b = 20; b >= 0; --b)
      for (int k = 2; k >= 0; --k) key = (key << 1) | ((x[k] >> b) & 1);
    return key;
  }

  // Finalise: gids in Morton order, coordinates, neighbours, index shapes.
  void Finalize() {
    int maxlev = 0;
    for (auto &a : leaves) maxlev = std::max(maxlev, a.level);
    std::vector<Loc> order(leaves.begin(), leaves.end());
    std::sort(order.begin(), order.end(), [&](const Loc &a, const Loc &b) {
      auto ma = Morton(a, maxlev), mb = Morton(b, maxlev);
      if (ma != mb) return ma < mb;
      return a.level < b.level;
    });
    blocks.clear();
    gid_of.clear();
    for (size_t g = 0; g < order.size(); ++g) {
      Block b;
      b.gid = (int)g;
      b.loc = order[g];
      BlockExtent(b.loc, b.xmin, b.xmax);
      gid_of[b.loc] = (int)g;
      blocks.push_back(b);
    }
    const int g = d.nghost;
    for (int k = 0; k < 3; ++k) {
      cellbounds.n[k] = active[k] ? d.bnx[k] : 1;
      cellbounds.ng[k] = active[k] ? g : 0;
      c_cellbounds.n[k] = active[k] ? std::max(1, d.bnx[k] / 2) : 1;
      c_cellbounds.ng[k] = active[k] ? g : 0;
    }
    for (auto &b : blocks) {
      Coords &c = b.coords;
      for (int k = 0; k < 3; ++k) {
        c.dx[k] = (b.xmax[k] - b.xmin[k]) / cellbounds.n[k];
        c.istart[k] = active[k] ? g : 0;
        c.xmin[k] = b.xmin[k] - c.istart[k] * c.dx[k];
      }
      c.area[0] = c.dx[1] * c.dx[2];
      c.area[1] = c.dx[0] * c.dx[2];
      c.area[2] = c.dx[0] * c.dx[1];
      c.vol = c.dx[0] * c.dx[1] * c.dx[2];
      // reference: uniform_cartesian.hpp:41-54 (coarsen = 2)
      Coords &cc = b.coarse_coords;
      cc = c;
      for (int k = 0; k < 3; ++k) cc.xmin[k] += c.istart[k] * c.dx[k] * (1 - 2);
      cc.dx[0] *= 2;
      cc.dx[1] *= (c.istart[1] > 0 ? 2 : 1);
      cc.dx[2] *= (c.istart[2] > 0 ? 2 : 1);
      cc.area[0] = cc.dx[1] * cc.dx[2];
      cc.area[1] = cc.dx[0] * cc.dx[2];
      cc.area[2] = cc.dx[0] * cc.dx[1];
      cc.vol = cc.dx[0] * cc.dx[1] * cc.dx[2];
    }
    for (auto &b : blocks) FindNeighbors(b);
  }

  // reference: logical_location.cpp:110-129
  bool IsNeighbor(const Loc &me, const Loc &in) const {
    int maxl = std::max(me.level, in.level);
    std::int64_t bs_in = 1LL << (maxl - in.level), bs_me = 1LL << (maxl - me.level);
    for (int k = 0; k < 3; ++k) {
      if (!active[k]) continue;
      std::int64_t low = me.l[k] * bs_me - 1, hi = low + bs_me + 1;
      std::int64_t low_in = in.l[k] * bs_in, hi_in = low_in + bs_in - 1;
      if (hi < low_in || low > hi_in) return false;
    }
    return true;
  }

  // reference: logical_location.cpp:100-108 (GetSameLevelOffsets)
  void SameLevelOffsets(const Loc &me, const Loc &nb, int *off) const {
    int sn = std::max(nb.level - me.level, 0), sm = std::max(me.level - nb.level, 0);
    for (int k = 0; k < 3; ++k) {
      std::int64_t a = nb.l[k], b = me.l[k];
      // arithmetic shifts of possibly negative (unwrapped) indices
      for (int s = 0; s < sn; ++s) a = floordiv2(a);
      for (int s = 0; s < sm; ++s) b = floordiv2(b);
      off[k] = (int)(a - b);
    }
  }

  void AddNeighbor(Block &b, const Loc &global, const Loc &origin) {
    NeighborBlock nb;
    nb.gid = gid_of.at(global);
    nb.loc = global;
    nb.origin_loc = origin;
    SameLevelOffsets(b.loc, origin, nb.off);
    if (origin.level < b.loc.level) b.has_coarser_neighbor = true;
    b.neighbors.push_back(nb);
  }

  // reference: tree.cpp:139-233 (leaf grid)
  void FindNeighbors(Block &b) {
    b.neighbors.clear();
    const Loc &me = b.loc;
    for (int o3 = (active[2] ? -1 : 0); o3 <= (active[2] ? 1 : 0); ++o3)
      for (int o2 = (active[1] ? -1 : 0); o2 <= (active[1] ? 1 : 0); ++o2)
        for (int o1 = -1; o1 <= 1; ++o1) {
          if (o1 == 0 && o2 == 0 && o3 == 0) continue;
          Loc n = me;
          n.l[0] += o1;
          n.l[1] += o2;
          n.l[2] += o3;
          Loc wn = n;
          if (!Wrap(wn)) continue;
          if (leaves.count(wn)) {
            AddNeighbor(b, wn, n);
          } else if (internal.count(wn)) {
            for (auto &dw : Daughters(wn)) {
              Loc dorig = dw;
              for (int k = 0; k < 3; ++k) dorig.l[k] += 2 * (n.l[k] - wn.l[k]);
              if (IsNeighbor(me, dorig)) AddNeighbor(b, dw, dorig);
            }
          } else {
            Loc pw = Parent(wn);
            if (leaves.count(pw)) {
              Loc porig = Parent(n);
              int so[3];
              SameLevelOffsets(me, porig, so);
              if (so[0] == o1 && so[1] == o2 && so[2] == o3) AddNeighbor(b, pw, porig);
            } else {
              throw std::runtime_error("mesh is not properly nested");
            }
          }
        }
  }
};

}  // namespace orc
