// Host-side mesh topology: block lattice, static refinement with 2:1 nesting, neighbour
// links and the index box of every boundary region. Everything here runs once at setup;
// its output is the flat region tables the device kernels walk.
//
// Behaviour follows the reference (paths relative to its root):
//   refinement and nesting      src/mesh/forest/tree.cpp:72-134, src/mesh/mesh.cpp:1054-1116
//   neighbour search            src/mesh/forest/tree.cpp:139-233, src/mesh/mesh-gmg.cpp:48-94
//   block extents               src/mesh/forest/logical_location.cpp:61-79, src/defs.hpp:98-101
//   coordinates                 src/coordinates/uniform_cartesian.hpp:30-58
//   region index boxes          src/bvals/comms/bnd_info.cpp:105-223 (CalcIndices)
// The root grid is one Cartesian lattice of base blocks with identity orientation, which is
// the reference's single-tree forest for power-of-two block counts.
#include <algorithm>
#include <cmath>

#include "pth_internal.hpp"

namespace pth {

namespace {
std::int64_t Half(std::int64_t v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }  // floor(v/2)
Lattice Up(const Lattice &a) {
  Lattice p;
  p.level = a.level - 1;
  for (int d = 0; d < 3; ++d) p.x[d] = Half(a.x[d]);
  return p;
}
// symmetrized logical coordinate in [-1/2, 1/2] of lattice index `index` (edge 0: left,
// 2: right) on a lattice of `range` blocks (reference IndexToSymmetrizedCoordinate)
double Symmetrized(std::int64_t index, int edge, std::int64_t range) {
  const std::int64_t a = index - range / 2, b = index - (range + 1) / 2;
  return static_cast<double>(a + b + edge) / (2.0 * range);
}
}  // namespace

std::string Topology::Init(const int *nx_, const int *bnx_, int g, const double *lo,
                           const double *hi, const int *per) {
  ng = g;
  ndim = 0;
  for (int d = 0; d < 3; ++d) {
    nx[d] = nx_[d];
    bnx[d] = bnx_[d];
    xmin[d] = lo[d];
    xmax[d] = hi[d];
    periodic[d] = per[d] != 0;
    bc[2 * d] = bc[2 * d + 1] = periodic[d] ? PTH_BC_PERIODIC : PTH_BC_OUTFLOW;
    active[d] = nx[d] > 1;
    if (active[d]) ndim = d + 1;
    if (bnx[d] < 1 || nx[d] % bnx[d] != 0) return "meshblock size must divide the mesh size";
    if (active[d] && bnx[d] < g) return "meshblock smaller than nghost";
    nbase[d] = nx[d] / bnx[d];
  }
  if (g < 1) return "nghost must be >= 1";
  leaf.clear();
  inner.clear();
  for (std::int64_t k = 0; k < nbase[2]; ++k)
    for (std::int64_t j = 0; j < nbase[1]; ++j)
      for (std::int64_t i = 0; i < nbase[0]; ++i) {
        Lattice a;
        a.x[0] = i;
        a.x[1] = j;
        a.x[2] = k;
        leaf.insert(a);
      }
  return "";
}

bool Topology::Fold(Lattice &a) const {
  for (int d = 0; d < 3; ++d) {
    const std::int64_t N = active[d] ? Span(d, a.level) : 1;
    if (a.x[d] >= 0 && a.x[d] < N) continue;
    if (!active[d] || !periodic[d]) return false;
    a.x[d] = ((a.x[d] % N) + N) % N;
  }
  return true;
}

std::vector<Lattice> Topology::Children(const Lattice &a) const {
  std::vector<Lattice> out;
  for (int k = 0; k < (active[2] ? 2 : 1); ++k)
    for (int j = 0; j < (active[1] ? 2 : 1); ++j)
      for (int i = 0; i < 2; ++i) {
        Lattice c;
        c.level = a.level + 1;
        c.x[0] = 2 * a.x[0] + i;
        c.x[1] = active[1] ? 2 * a.x[1] + j : 0;
        c.x[2] = active[2] ? 2 * a.x[2] + k : 0;
        out.push_back(c);
      }
  return out;
}

// Refine one leaf, then refine the leaves on the parent level that touch the corner of the
// parent nearest to it, so every child keeps neighbours at most one level coarser. Returns the
// number of blocks added (children minus the split leaf, over the whole cascade).
int Topology::Split(const Lattice &a) {
  if (!leaf.count(a)) return 0;
  leaf.erase(a);
  inner.insert(a);
  const std::vector<Lattice> kids = Children(a);
  for (const auto &c : kids) leaf.insert(c);
  multilevel = true;
  int added = static_cast<int>(kids.size()) - 1;
  if (a.level == 0) return added;
  const Lattice p = Up(a);
  int corner[3];
  for (int d = 0; d < 3; ++d) corner[d] = static_cast<int>(a.x[d] - 2 * p.x[d]);
  for (int k = 0; k < (active[2] ? 2 : 1); ++k)
    for (int j = 0; j < (active[1] ? 2 : 1); ++j)
      for (int i = 0; i < 2; ++i) {
        Lattice q = p;
        q.x[0] += i + corner[0] - 1;
        if (active[1]) q.x[1] += j + corner[1] - 1;
        if (active[2]) q.x[2] += k + corner[2] - 1;
        if (Fold(q)) added += Split(q);
      }
  return added;
}

int Topology::Refine(const Lattice &a) { return Split(a); }

int Topology::Derefine(const Lattice &parent) {
  const std::vector<Lattice> kids = Children(parent);
  for (const Lattice &c : kids) {
    if (!leaf.count(c)) return 0;
    for (int k = active[2] ? -1 : 0; k <= (active[2] ? 1 : 0); ++k)
      for (int j = active[1] ? -1 : 0; j <= (active[1] ? 1 : 0); ++j)
        for (int i = -1; i <= 1; ++i) {
          Lattice q = c;
          q.x[0] += i;
          q.x[1] += j;
          q.x[2] += k;
          if (Fold(q) && inner.count(q)) return 0;
        }
  }
  for (const Lattice &c : kids) leaf.erase(c);
  inner.erase(parent);
  leaf.insert(parent);
  return static_cast<int>(kids.size()) - 1;
}

int Topology::RootLevel() const {
  std::int64_t div = std::int64_t(1) << 40;
  for (int d = 0; d < 3; ++d) {
    if (!active[d]) {
      div = 1;
      continue;
    }
    std::int64_t p = 1;
    while (nbase[d] % (2 * p) == 0) p *= 2;
    div = std::min(div, p);
  }
  int lev = 0;
  while ((std::int64_t(1) << (lev + 1)) <= div) ++lev;
  return lev;
}

bool Topology::Nested() const {
  // every position next to a leaf is a leaf, an inner node whose children touching the leaf
  // are leaves, or lies inside a leaf one level coarser
  for (const Lattice &a : leaf)
    for (int k = active[2] ? -1 : 0; k <= (active[2] ? 1 : 0); ++k)
      for (int j = active[1] ? -1 : 0; j <= (active[1] ? 1 : 0); ++j)
        for (int i = -1; i <= 1; ++i) {
          Lattice seen = a;
          seen.x[0] += i;
          seen.x[1] += j;
          seen.x[2] += k;
          Lattice q = seen;
          if (!Fold(q) || leaf.count(q)) continue;
          if (inner.count(q)) {
            for (const Lattice &c : Children(seen)) {
              bool touches = true;
              for (int d = 0; d < 3; ++d)
                if (active[d] && (c.x[d] < 2 * a.x[d] - 1 || c.x[d] > 2 * a.x[d] + 2))
                  touches = false;
              Lattice cf = c;
              if (touches && Fold(cf) && !leaf.count(cf)) return false;
            }
          } else if (q.level == 0 || !leaf.count(Up(q))) {
            return false;
          }
        }
  return true;
}

// Make `a` a block of the mesh by refining its ancestors from the coarsest leaf down.
void Topology::Ensure(const Lattice &a) {
  if (leaf.count(a) || inner.count(a)) return;
  std::vector<Lattice> chain;
  Lattice p = Up(a);
  for (int lev = a.level - 1; lev >= 0; --lev) {
    chain.push_back(p);
    if (leaf.count(p)) break;
    p = Up(p);
  }
  for (auto it = chain.rbegin(); it != chain.rend(); ++it) Split(*it);
}

double Topology::Position(int d, std::int64_t index, int edge, int level) const {
  const double u = Symmetrized(index, edge, Span(d, level));
  return 0.5 * (xmin[d] + xmax[d]) + (u * xmax[d] - u * xmin[d]);
}

void Topology::RefineBox(const double box[6], int level) {
  std::int64_t lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  for (int d = 0; d < 3; ++d) {
    if (!active[d]) continue;
    const std::int64_t tot = Span(d, level);
    std::int64_t a = 0;
    while (a < tot && !(Position(d, a + 1, 0, level) > box[2 * d])) ++a;
    std::int64_t b = a;
    while (b < tot && !(Position(d, b + 1, 0, level) >= box[2 * d + 1])) ++b;
    if (a % 2 == 1) --a;
    if (b % 2 == 0) ++b;
    lo[d] = a;
    hi[d] = b;
  }
  for (std::int64_t k = lo[2]; k < hi[2]; k += 2)
    for (std::int64_t j = lo[1]; j < hi[1]; j += 2)
      for (std::int64_t i = lo[0]; i < hi[0]; i += 2) {
        Lattice a;
        a.level = level;
        a.x[0] = i;
        a.x[1] = j;
        a.x[2] = k;
        Ensure(a);
      }
}

void Topology::RefineSphere(const double c[3], double r, int level) {
  const std::int64_t N0 = Span(0, level), N1 = active[1] ? Span(1, level) : 1,
                     N2 = active[2] ? Span(2, level) : 1;
  for (std::int64_t k = 0; k < N2; ++k)
    for (std::int64_t j = 0; j < N1; ++j)
      for (std::int64_t i = 0; i < N0; ++i) {
        const std::int64_t ix[3] = {i, j, k};
        double d2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          if (!active[d]) continue;
          const double lo = Position(d, ix[d], 0, level), hi = Position(d, ix[d], 2, level);
          const double p = std::min(std::max(c[d], lo), hi);
          d2 += (p - c[d]) * (p - c[d]);
        }
        if (d2 < r * r) {
          Lattice a;
          a.level = level;
          a.x[0] = i;
          a.x[1] = j;
          a.x[2] = k;
          Ensure(a);
        }
      }
}

std::string Topology::Finalize() {
  int top = 0;
  for (const auto &a : leaf) top = std::max(top, a.level);
  // Morton order of the leaves (interleaved bits at the finest level)
  auto key = [&](const Lattice &a) {
    std::uint64_t k = 0;
    for (int bit = 20; bit >= 0; --bit)
      for (int d = 2; d >= 0; --d)
        k = (k << 1) | ((static_cast<std::uint64_t>(a.x[d] << (top - a.level)) >> bit) & 1u);
    return k;
  };
  std::vector<Lattice> order(leaf.begin(), leaf.end());
  std::sort(order.begin(), order.end(), [&](const Lattice &a, const Lattice &b) {
    const auto ka = key(a), kb = key(b);
    return ka != kb ? ka < kb : a.level < b.level;
  });
  for (int d = 0; d < 3; ++d) {
    fine.n[d] = active[d] ? bnx[d] : 1;
    fine.g[d] = active[d] ? ng : 0;
    coarse.n[d] = active[d] ? std::max(1, bnx[d] / 2) : 1;
    coarse.g[d] = active[d] ? ng : 0;
  }
  if (multilevel) {
    if (ng % 2 != 0) return "mesh refinement needs an even nghost";
    for (int d = 0; d < 3; ++d)
      if (active[d] && bnx[d] % 2 != 0) return "mesh refinement needs even meshblock sizes";
  }
  blocks.assign(order.size(), Block());
  id.clear();
  for (size_t gid = 0; gid < order.size(); ++gid) {
    Block &b = blocks[gid];
    b.at = order[gid];
    id[b.at] = static_cast<int>(gid);
    for (int d = 0; d < 3; ++d) {
      if (active[d]) {
        b.xmin[d] = Position(d, b.at.x[d], 0, b.at.level);
        b.xmax[d] = Position(d, b.at.x[d], 2, b.at.level);
      } else {
        b.xmin[d] = xmin[d];
        b.xmax[d] = xmax[d];
      }
      b.dx[d] = (b.xmax[d] - b.xmin[d]) / fine.n[d];
      const int ist = active[d] ? ng : 0;
      b.x0[d] = b.xmin[d] - ist * b.dx[d];
      b.cx0[d] = b.x0[d] + ist * b.dx[d] * (1 - 2);
    }
    b.cdx[0] = b.dx[0] * 2;
    b.cdx[1] = b.dx[1] * (active[1] ? 2 : 1);
    b.cdx[2] = b.dx[2] * (active[2] ? 2 : 1);
    b.area[0] = b.dx[1] * b.dx[2];
    b.area[1] = b.dx[0] * b.dx[2];
    b.area[2] = b.dx[0] * b.dx[1];
    b.vol = b.dx[0] * b.dx[1] * b.dx[2];
  }
  for (size_t gid = 0; gid < blocks.size(); ++gid) Neighbors(static_cast<int>(gid));
  return "";
}

void Topology::Neighbors(int gid) {
  Block &b = blocks[gid];
  b.links.clear();
  auto add = [&](const Lattice &folded, const Lattice &seen) {
    Link l;
    l.gid = id.at(folded);
    l.level = seen.level;
    for (int d = 0; d < 3; ++d) {
      l.seen[d] = seen.x[d];
      // offset direction on the coarser of the two levels
      std::int64_t a = seen.x[d], c = b.at.x[d];
      for (int s = b.at.level; s < seen.level; ++s) a = Half(a);
      for (int s = seen.level; s < b.at.level; ++s) c = Half(c);
      l.off[d] = static_cast<int>(a - c);
    }
    if (seen.level < b.at.level) b.coarser_neighbor = true;
    b.links.push_back(l);
  };
  for (int o2 = active[2] ? -1 : 0; o2 <= (active[2] ? 1 : 0); ++o2)
    for (int o1 = active[1] ? -1 : 0; o1 <= (active[1] ? 1 : 0); ++o1)
      for (int o0 = -1; o0 <= 1; ++o0) {
        if (o0 == 0 && o1 == 0 && o2 == 0) continue;
        Lattice seen = b.at;
        seen.x[0] += o0;
        seen.x[1] += o1;
        seen.x[2] += o2;
        Lattice f = seen;
        if (!Fold(f)) continue;
        if (leaf.count(f)) {
          add(f, seen);
        } else if (inner.count(f)) {
          // finer neighbours: children of the same-level position that touch this block
          for (const auto &c : Children(f)) {
            Lattice cs = c;
            bool touches = true;
            for (int d = 0; d < 3; ++d) {
              cs.x[d] += 2 * (seen.x[d] - f.x[d]);
              if (!active[d]) continue;
              const std::int64_t lo = 2 * b.at.x[d] - 1, hi = 2 * b.at.x[d] + 2;
              if (cs.x[d] < lo || cs.x[d] > hi) touches = false;
            }
            if (touches) add(c, cs);
          }
        } else {
          // coarser neighbour, listed once under the direction it lies in
          const Lattice pf = Up(f), ps = Up(seen);
          if (!leaf.count(pf)) continue;  // unreachable for a properly nested mesh
          std::int64_t me[3] = {Half(b.at.x[0]), Half(b.at.x[1]), Half(b.at.x[2])};
          if (ps.x[0] - me[0] == o0 && ps.x[1] - me[1] == o1 && ps.x[2] - me[2] == o2)
            add(pf, ps);
        }
      }
}

// Index box of the region exchanged with neighbour `l` (reference CalcIndices for a cell-
// centred, non-flux field). Boxes of a receiving side are the ghost strips, of a sending
// side the interior strips; the coarse index space is used for prolongation/restriction
// and whenever the neighbour is coarser.
Box Topology::Indices(const Block &b, const Link &l, Side side, bool prores) const {
  const Extent &sh = (prores || l.level < b.at.level) ? coarse : fine;
  const int factor = l.level > b.at.level ? 2 : 1;
  const int in_off = side == Side::kSendBoundary ? ng : 0;
  int ex_off = side == Side::kRecvBoundary ? ng : 0;
  if (prores) ex_off /= 2;
  Box out;
  for (int d = 0; d < 3; ++d) {
    const int lo = sh.lo(d), hi = sh.hi(d);
    int s, e;
    if (l.off[d] > 0) {
      s = hi - in_off + 1;
      e = hi + ex_off;
    } else if (l.off[d] < 0) {
      s = lo - ex_off;
      e = lo + in_off - 1;
    } else {
      s = lo;
      e = hi;
      if (active[d] && b.at.level < l.level) {
        // finer neighbour covers half of this block; send g extra cells for its prolongation
        const int extra = (hi - lo + 1) - (active[d] ? bnx[d] / factor : 1);
        if (l.seen[d] % 2 == 1) s += extra - in_off;
        if (l.seen[d] % 2 == 0) e -= extra - in_off;
        if (side == Side::kSendInterior && !prores) {
          s -= ng;
          e += ng;
        }
      }
      if (active[d] && b.at.level > l.level) {
        // coarser neighbour sent extra cells on the side this block's parent extends to
        if (b.at.x[d] % 2 == 1) s -= ex_off;
        if (b.at.x[d] % 2 == 0) e += ex_off;
        if (side == Side::kRecvInterior && !prores) {
          s -= ng;
          e += ng;
        }
      }
      if (prores && active[d] && side == Side::kRecvInterior) {
        s -= ng / 2;
        e += ng / 2;
      }
    }
    out.lo[d] = s;
    out.hi[d] = e;
  }
  return out;
}

}  // namespace pth
