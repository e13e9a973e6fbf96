// ORACLE — TEST INFRASTRUCTURE ONLY.
// CPU restatement of the reference's adaptive remesh: refinement tagging, the per-block
// refinement flags, the tree update and the movement of cell data onto the new block list.
// Only tests/, the smoke check and the CPU-baseline leg of bench.py may load this code; the
// product library never links it.
//
// Restated from (paths relative to the reference root):
//   src/amr_criteria/refinement_package.cpp:51-91     CheckAllRefinement
//   src/amr_criteria/refinement_package.cpp:93-151    FirstDerivative / SecondDerivative
//   src/amr_criteria/amr_criteria.cpp:27-68,88-106    criterion parameters and operator()
//   src/interface/state_descriptor.hpp:338-341         package default: derefine
//   src/mesh/mesh_refinement.cpp:55-58,85-117          deref_count threshold, SetRefinement
//   src/mesh/mesh-amr_loadbalance.cpp:324-351          LoadBalancingAndAdaptiveMeshRefinement
//   src/mesh/mesh-amr_loadbalance.cpp:496-626          UpdateMeshBlockTree
//   src/mesh/mesh-amr_loadbalance.cpp:663-760          new block list, restriction of f2c blocks
//   src/mesh/mesh-amr_loadbalance.cpp:83-140,178-240   TryRecvCoarseToFine / TryRecvFineToCoarse
//   src/mesh/mesh-amr_loadbalance.cpp:864-894,995-1003 interior prolongation, CommunicateBoundaries
//   src/bvals/comms/bnd_info.cpp:355-385,198-203       interior restrict / prolongate index ranges
//   src/driver/driver.cpp:120-131,194-235               remesh in the evolution loop, time step
//   src/config.hpp.in:83                                TINY_NUMBER
#pragma once
#include <algorithm>
#include <memory>
#include <stdexcept>
#include <vector>

#include "orc_burgers.hpp"
#include "orc_bvals.hpp"
#include "orc_mesh.hpp"

namespace orc {

constexpr double kTinyNumber = 1.0e-20;

// One <parthenon/refinementN> block: method 1 = derivative_order_1, 2 = derivative_order_2.
struct Criterion {
  int method = 1;
  int comp = 0;  // vector_i
  double refine_tol = 0.5, derefine_tol = 0.05;
  int max_level = 1;  // compared with the reference's logical level (root_level + level)
};

// Largest derivative measure of one component slab over the block interior.
inline double DerivativeMax(const Mesh &m, const double *q, int method) {
  const Shape &s = m.cellbounds;
  const int ndim = 1 + (s.e(1) > s.s(1)) + (s.e(2) > s.s(2));
  const std::int64_t sj = s.ent(0), sk = (std::int64_t)s.ent(0) * s.ent(1);
  double maxd = 0.0;
  for (int k = s.s(2); k <= s.e(2); ++k)
    for (int j = s.s(1); j <= s.e(1); ++j)
      for (int i = s.s(0); i <= s.e(0); ++i) {
        const std::int64_t c = k * sk + j * sj + i;
        const std::int64_t st[3] = {1, sj, sk};
        for (int d = 0; d < ndim; ++d) {
          double v;
          if (method == 1) {
            const double scale = std::abs(q[c]);
            v = 0.5 * std::abs((q[c + st[d]] - q[c - st[d]])) / (scale + kTinyNumber);
          } else {
            const double aqt = std::abs(q[c]) + kTinyNumber;
            const double qavg = 0.5 * (q[c + st[d]] + q[c - st[d]]);
            v = std::abs(qavg - q[c]) / (std::abs(qavg) + aqt);
          }
          maxd = (v > maxd ? v : maxd);
        }
      }
  return maxd;
}

inline int TagOf(double maxd, const Criterion &c) {
  if (maxd > c.refine_tol) return 1;
  if (maxd < c.derefine_tol) return -1;
  return 0;
}

// Refinement state of an adaptive mesh: criteria and the per-block MeshRefinement members.
struct Adaptive {
  int numlevel = 1, deref_threshold = 10, root_level = 0, max_level = 0;
  std::vector<Criterion> crit;
  std::vector<int> flag, count;  // refine_flag_, deref_count_ per block (gid order)

  void Init(const Mesh &m, int numlevel_, int threshold) {
    numlevel = numlevel_;
    deref_threshold = threshold;
    root_level = m.RootLevel();
    max_level = numlevel + root_level - 1;
    flag.assign(m.blocks.size(), 0);
    count.assign(m.blocks.size(), 0);
  }

  // AMRCriteria constructor: the criterion's max_level defaults to and is clamped at numlevel
  void AddCriterion(Criterion c) {
    if (c.max_level <= 0 || c.max_level > numlevel) c.max_level = numlevel;
    crit.push_back(c);
  }

  // CheckAllRefinement for block b of field data f ([block][comp][k][j][i])
  int CheckAll(const Mesh &m, const Field &f, int b) const {
    int delta = -1;  // a package without a block check recommends derefinement
    for (const Criterion &c : crit) {
      const double *q = f.data.data() + ((std::int64_t)b * f.ncomp + c.comp) * m.cellbounds.size();
      int t = TagOf(DerivativeMax(m, q, c.method), c);
      if (t == 1 && root_level + m.blocks[b].loc.level >= c.max_level) t = 0;
      delta = std::max(delta, t);
      if (delta == 1) return 1;
    }
    return delta;
  }

  // MeshRefinement::SetRefinement
  void SetRefinement(const Mesh &m, int b, int tag) {
    const int aret = std::max(-1, tag);
    const int level = root_level + m.blocks[b].loc.level;
    if (aret == 0) flag[b] = 0;
    if (aret >= 0) count[b] = 0;
    if (aret > 0) {
      flag[b] = level == max_level ? 0 : 1;
    } else if (aret < 0) {
      if (level == root_level) {
        flag[b] = 0;
        count[b] = 0;
      } else {
        count[b]++;
        int finer = 0;
        for (const NeighborBlock &nb : m.blocks[b].neighbors)
          if (nb.loc.level > m.blocks[b].loc.level) finer++;
        if (finer > 0)
          flag[b] = 0;
        else
          flag[b] = count[b] >= deref_threshold ? -1 : 0;
      }
    }
  }

  // Refinement::Tag over every block
  void Tag(const Mesh &m, const Field &f) {
    for (int b = 0; b < (int)m.blocks.size(); ++b) SetRefinement(m, b, CheckAll(m, f, b));
  }

  // Mesh::UpdateMeshBlockTree applied to `t`, a copy of m's tree. Derefinement candidates are
  // parents whose nleaf daughters appear consecutively (gid order) among the blocks flagged -1;
  // the candidate list is sorted coarsest-last over all entries but the last one, as the
  // reference's sort bound does.
  void UpdateTree(const Mesh &m, Mesh &t, int &nnew, int &ndel) const {
    nnew = ndel = 0;
    const int nleaf = 1 << m.ndim;
    std::vector<Loc> lref, lderef;
    for (const Block &b : m.blocks) {
      if (flag[b.gid] == 1) lref.push_back(b.loc);
      if (flag[b.gid] == -1) lderef.push_back(b.loc);
    }
    if (lref.empty() && (int)lderef.size() < nleaf) return;
    std::vector<Loc> clderef;
    if ((int)lderef.size() >= nleaf) {
      const int lj = m.active[1] ? 1 : 0, lk = m.active[2] ? 1 : 0;
      for (size_t n = 0; n < lderef.size(); ++n) {
        const Loc &a = lderef[n];
        if ((a.l[0] & 1) || (a.l[1] & 1) || (a.l[2] & 1)) continue;
        size_t r = n;
        int rr = 0;
        for (int k = 0; k <= lk; ++k)
          for (int j = 0; j <= lj; ++j)
            for (int i = 0; i <= 1; ++i) {
              if (r < lderef.size()) {
                const Loc &c = lderef[r];
                if (a.l[0] + i == c.l[0] && a.l[1] + j == c.l[1] && a.l[2] + k == c.l[2] &&
                    a.level == c.level)
                  rr++;
                r++;
              }
            }
        if (rr == nleaf) clderef.push_back(Parent(a));
      }
    }
    if (clderef.size() > 1)
      std::stable_sort(clderef.begin(), clderef.end() - 1,
                       [](const Loc &x, const Loc &y) { return x.level > y.level; });
    t.leaves = m.leaves;
    t.internal = m.internal;
    for (const Loc &a : lref) nnew += t.Refine(a);
    for (const Loc &p : clderef) ndel += t.Derefine(p);
  }
};

// How a block of the new list gets its data (RedistributeAndRefineMeshBlocks).
struct Origin {
  int kind;     // 0 same level (moved), 1 coarse to fine, 2 fine to coarse
  int old[8];   // old gid(s): one for kinds 0 and 1, the nleaf daughters for kind 2
  int ox[8][3]; // kind 1: position of the new block in its old parent; kind 2: of each daughter
  int n = 1;
};

inline std::vector<Origin> Origins(const Mesh &om, const Mesh &nm) {
  std::vector<Origin> out(nm.blocks.size());
  for (const Block &b : nm.blocks) {
    Origin &o = out[b.gid];
    auto it = om.gid_of.find(b.loc);
    if (it != om.gid_of.end()) {
      o.kind = 0;
      o.old[0] = it->second;
      continue;
    }
    if (b.loc.level > 0) {
      auto ip = om.gid_of.find(Parent(b.loc));
      if (ip != om.gid_of.end()) {
        o.kind = 1;
        o.old[0] = ip->second;
        for (int d = 0; d < 3; ++d) o.ox[0][d] = (int)(b.loc.l[d] & 1);
        continue;
      }
    }
    o.kind = 2;
    o.n = 0;
    for (const Loc &c : nm.Daughters(b.loc)) {
      auto ic = om.gid_of.find(c);
      if (ic == om.gid_of.end()) throw std::runtime_error("remesh changed a block by more than one level");
      o.old[o.n] = ic->second;
      for (int d = 0; d < 3; ++d) o.ox[o.n][d] = (int)(c.l[d] & 1);
      o.n++;
    }
  }
  return out;
}

// Cell data of one field moved from the old to the new block list. Derefined blocks restrict
// their interior into their own coarse buffer first; the coarse parent copies those octants
// into its interior. A refined block receives the whole coarse-shaped brick of its old parent
// around its octant (ghosts included) into its coarse buffer and prolongates the coarse
// interior extended by nghost/2 cells, which covers its whole brick. Same-level blocks move
// whole. New arrays start at zero like a freshly allocated block.
inline void TransferField(const Mesh &om, Field &of, const Mesh &nm, Field &nf) {
  const Shape &fs = om.cellbounds, &cs = om.c_cellbounds;
  const int V = of.ncomp, g = om.d.nghost;
  nf.ncomp = V;
  nf.data.assign(nm.blocks.size() * V * fs.size(), 0.0);
  nf.coarse.assign(nm.blocks.size() * V * cs.size(), 0.0);
  if (of.coarse.size() != om.blocks.size() * V * cs.size())
    of.coarse.assign(om.blocks.size() * V * cs.size(), 0.0);
  const std::vector<Origin> orig = Origins(om, nm);
  // restriction of every derefined block (interior of its coarse buffer)
  for (const Origin &o : orig) {
    if (o.kind != 2) continue;
    for (int q = 0; q < o.n; ++q)
      for (int c = 0; c < V; ++c)
        for (int k = cs.s(2); k <= cs.e(2); ++k)
          for (int j = cs.s(1); j <= cs.e(1); ++j)
            for (int i = cs.s(0); i <= cs.e(0); ++i)
              RestrictCell(om, om.blocks[o.old[q]], of, c, k, j, i);
  }
  for (const Block &b : nm.blocks) {
    const Origin &o = orig[b.gid];
    double *dst = BlockSlab(nf, fs, b.gid, false);
    if (o.kind == 0) {
      const double *src = BlockSlab(of, fs, o.old[0], false);
      std::copy(src, src + V * fs.size(), dst);
    } else if (o.kind == 1) {
      double *src = BlockSlab(of, fs, o.old[0], false);
      double *cdst = BlockSlab(nf, cs, b.gid, true);
      int off[3];
      for (int d = 0; d < 3; ++d) off[d] = o.ox[0][d] ? fs.n[d] / 2 : 0;
      for (int c = 0; c < V; ++c)
        for (int k = 0; k < cs.ent(2); ++k)
          for (int j = 0; j < cs.ent(1); ++j)
            for (int i = 0; i < cs.ent(0); ++i)
              At(cdst, cs, c, k, j, i) = At(src, fs, c, k + off[2], j + off[1], i + off[0]);
    } else {
      for (int q = 0; q < o.n; ++q) {
        double *csrc = BlockSlab(of, cs, o.old[q], true);
        int off[3];
        for (int d = 0; d < 3; ++d) off[d] = o.ox[q][d] ? cs.n[d] : 0;
        for (int c = 0; c < V; ++c)
          for (int k = cs.s(2); k <= cs.e(2); ++k)
            for (int j = cs.s(1); j <= cs.e(1); ++j)
              for (int i = cs.s(0); i <= cs.e(0); ++i)
                At(dst, fs, c, k + off[2], j + off[1], i + off[0]) = At(csrc, cs, c, k, j, i);
      }
    }
  }
  // interior prolongation of every refined block
  int lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = nm.active[d] ? cs.s(d) - g / 2 : 0;
    hi[d] = nm.active[d] ? cs.e(d) + g / 2 : 0;
  }
  for (const Block &b : nm.blocks) {
    if (orig[b.gid].kind != 1) continue;
    for (int c = 0; c < V; ++c)
      for (int k = lo[2]; k <= hi[2]; ++k)
        for (int j = lo[1]; j <= hi[1]; ++j)
          for (int i = lo[0]; i <= hi[0]; ++i) ProlongateCell(nm, b, nf, c, k, j, i, true);
  }
}

// The new mesh of an updated tree (same description, multilevel storage).
inline std::unique_ptr<Mesh> MeshFromTree(const Mesh &tree) {
  std::unique_ptr<Mesh> nm(new Mesh(tree.d));
  nm->leaves = tree.leaves;
  nm->internal = tree.internal;
  nm->multilevel = true;
  nm->Finalize();
  return nm;
}

// Per-block flags carried to the new list: moved blocks keep theirs, new blocks start at 0.
inline void CarryFlags(const std::vector<Origin> &orig, Adaptive &a) {
  std::vector<int> f(orig.size(), 0), c(orig.size(), 0);
  for (size_t n = 0; n < orig.size(); ++n)
    if (orig[n].kind == 0) {
      f[n] = a.flag[orig[n].old[0]];
      c[n] = a.count[orig[n].old[0]];
    }
  a.flag.swap(f);
  a.count.swap(c);
}

// The Burgers benchmark on an adaptive mesh (EvolutionDriver::Execute): after each cycle, tag
// the base container, update the tree and, when it changed, move the data, refill ghosts,
// FillDerived and redo the time step from the new blocks (InitializeBlockTimeSteps, then
// SetGlobalTimeStep from the step size the cycle used). `mesh` owns the current mesh.
inline bool AdaptBurgers(std::unique_ptr<Mesh> &mesh, Burgers &B, Adaptive &amr) {
  std::unique_ptr<Mesh> tree(new Mesh(mesh->d));
  int nnew = 0, ndel = 0;
  amr.UpdateTree(*mesh, *tree, nnew, ndel);
  if (nnew == 0 && ndel == 0) return false;
  std::unique_ptr<Mesh> nm = MeshFromTree(*tree);
  const std::vector<Origin> orig = Origins(*mesh, *nm);
  Field nf;
  TransferField(*mesh, B.base.U, *nm, nf);
  Burgers nb;
  nb.Init(*nm, B.V - 3, B.cfl, B.recon);
  nb.base.U.data.swap(nf.data);
  nb.base.U.coarse.swap(nf.coarse);
  nb.dt = B.dt;
  nb.time = B.time;
  nb.ncycle = B.ncycle;
  nb.flux_correction = B.flux_correction;
  CarryFlags(orig, amr);
  B = std::move(nb);
  mesh = std::move(nm);
  B.m = mesh.get();
  B.BuildExchanges();
  B.DoExchange(B.base, B.ex_base, B.fb);
  B.CalculateDerived(B.base);
  return true;
}

// One cycle of the adaptive Burgers run; returns true when the mesh was changed after it.
inline bool AdaptiveCycle(std::unique_ptr<Mesh> &mesh, Burgers &B, Adaptive &amr) {
  const double dt_used = B.dt;
  B.Cycle();
  amr.Tag(*mesh, B.base.U);
  if (!AdaptBurgers(mesh, B, amr)) return false;
  B.dt = dt_used;
  B.SetGlobalTimeStep(B.EstimateTimestep(B.base));
  return true;
}

}  // namespace orc
