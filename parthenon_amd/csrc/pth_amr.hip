// Adaptive mesh refinement: the derivative criteria of Refinement::Tag on the device, the
// per-block refinement flags, the tree update, and the remesh that rebuilds the block list and
// moves every field's data onto it with three table-walking launches (restriction of derefined
// blocks, box copies, interior prolongation of refined blocks).
//
// Behaviour follows the reference (paths relative to its root):
//   derivative criteria         src/amr_criteria/refinement_package.cpp:51-151,
//                               src/amr_criteria/amr_criteria.cpp:27-68, 88-106
//   refinement flags            src/mesh/mesh_refinement.cpp:55-58, 85-117
//   tree update                 src/mesh/mesh-amr_loadbalance.cpp:496-626,
//                               src/mesh/forest/tree.cpp:96-137, 228-275
//   data movement               src/mesh/mesh-amr_loadbalance.cpp:83-140, 178-240, 739-900
//   interior index ranges       src/bvals/comms/bnd_info.cpp:105-223, 355-385
// The reductions are maxima, which are exact, so the tags equal the CPU restatement's bit for
// bit; restriction and prolongation reuse the boundary kernels' arithmetic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>

#include "pth_burgers.hpp"
#include "pth_internal.hpp"

namespace pth {

namespace {

constexpr int kTagWG = 256;
constexpr double kTinyNumber = 1.0e-20;  // reference TINY_NUMBER (config.hpp.in:83)

struct TagArgs {
  const double *q;  // field data, component 0 of block 0
  int ncomp, comp;
  int e[3];         // brick extents
  int lo[3], n[3];  // interior start and size
  int ndim;         // 1 + (n[1] > 1) + (n[2] > 1)
  int method;       // 1 first, 2 second derivative
};

__device__ __forceinline__ double MaxOf(double a, double b) { return b > a ? b : a; }

// One workgroup per block: the largest derivative measure over the block interior. NaN
// measures never replace the running maximum, as in the reference's `d > maxd ? d : maxd`.
__global__ void __launch_bounds__(kTagWG)
    DerivativeMaxKernel(TagArgs a, double *__restrict__ out, int stride, int slot) {
  const int b = blockIdx.x;
  const std::int64_t cells = static_cast<std::int64_t>(a.e[0]) * a.e[1] * a.e[2];
  const double *q = a.q + (static_cast<std::int64_t>(b) * a.ncomp + a.comp) * cells;
  const int st[3] = {1, a.e[0], a.e[0] * a.e[1]};
  const int n01 = a.n[0] * a.n[1], total = n01 * a.n[2];
  double m = 0.0;
  for (int t = threadIdx.x; t < total; t += kTagWG) {
    const int i = a.lo[0] + t % a.n[0];
    const int j = a.lo[1] + (t % n01) / a.n[0];
    const int k = a.lo[2] + t / n01;
    const int c = k * st[2] + j * st[1] + i;
    const double qc = q[c];
    for (int d = 0; d < a.ndim; ++d) {
      const double qm = q[c - st[d]], qp = q[c + st[d]];
      double v;
      if (a.method == 1) {
        v = 0.5 * fabs(qp - qm) / (fabs(qc) + kTinyNumber);
      } else {
        const double qavg = 0.5 * (qp + qm);
        v = fabs(qavg - qc) / (fabs(qavg) + (fabs(qc) + kTinyNumber));
      }
      m = v > m ? v : m;
    }
  }
  for (int off = 32; off > 0; off >>= 1) m = MaxOf(m, __shfl_xor(m, off, 64));
  __shared__ double w[kTagWG / 64];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = w[0];
    for (int q2 = 1; q2 < kTagWG / 64; ++q2) r = MaxOf(r, w[q2]);
    out[static_cast<std::int64_t>(b) * stride + slot] = r;
  }
}

Lattice Parent(const Lattice &a) {
  Lattice p;
  p.level = a.level - 1;
  for (int d = 0; d < 3; ++d) p.x[d] = a.x[d] >= 0 ? a.x[d] / 2 : -((1 - a.x[d]) / 2);
  return p;
}

// A region record for the remesh tables: box `lo`/`n` in a's index space.
Region MakeBox(double *a, double *b, const int lo[3], const int n[3], int nc, int gid,
               const Extent &ea, const Extent &eb) {
  Region r;
  std::memset(&r, 0, sizeof(r));
  r.a = a;
  r.b = b;
  r.nc = nc;
  r.gid = gid;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = lo[d];
    r.n[d] = n[d];
    r.magic[d] = Magic(n[d]);
  }
  r.astride[0] = ea.ent(0);
  r.astride[1] = ea.ent(0) * ea.ent(1);
  r.astride[2] = static_cast<int>(ea.cells());
  r.bstride[0] = eb.ent(0);
  r.bstride[1] = eb.ent(0) * eb.ent(1);
  r.bstride[2] = static_cast<int>(eb.cells());
  return r;
}

// How a block of the new list gets its data.
struct Origin {
  int kind = 0;  // 0 moved (same level), 1 coarse to fine, 2 fine to coarse
  int n = 1;
  int old[8];
  int ox[8][3];  // kind 1: octant of the new block in its old parent; kind 2: of each daughter
};

}  // namespace

void Mesh::SetRefinement(int b, int tag) {
  const int aret = std::max(-1, tag);
  const int root = topo.RootLevel();
  const int level = root + topo.blocks[b].at.level;
  const int max_level = amr.numlevel + root - 1;
  if (aret == 0) amr.flag[b] = 0;
  if (aret >= 0) amr.count[b] = 0;
  if (aret > 0) {
    amr.flag[b] = level == max_level ? 0 : 1;
  } else if (aret < 0) {
    if (level == root) {
      amr.flag[b] = 0;
      amr.count[b] = 0;
    } else {
      amr.count[b]++;
      int finer = 0;
      for (const Link &l : topo.blocks[b].links)
        if (l.level > topo.blocks[b].at.level) finer++;
      if (finer > 0)
        amr.flag[b] = 0;
      else
        amr.flag[b] = amr.count[b] >= amr.deref_threshold ? -1 : 0;
    }
  }
}

pth_status Mesh::Tag(MeshData &md) {
  if (!amr.on) {
    SetError("the mesh is not adaptive (pth_mesh_set_adaptive)");
    return PTH_FAIL;
  }
  const int nb = static_cast<int>(topo.blocks.size());
  const int nc = static_cast<int>(amr.crit.size());
  std::vector<double> maxd(static_cast<size_t>(nb) * std::max(nc, 1), 0.0);
  if (nc > 0) {
    const std::int64_t need = static_cast<std::int64_t>(nb) * nc;
    if (amr.maxd_n < need) {
      if (amr.maxd) (void)hipFree(amr.maxd);
      amr.maxd = nullptr;
      amr.maxd_n = 0;
      PTH_HIP(hipMalloc(&amr.maxd, sizeof(double) * need));
      amr.maxd_n = need;
    }
    for (int q = 0; q < nc; ++q) {
      const AmrCriterion &c = amr.crit[q];
      const FieldStore &st = md.store[c.field];
      if (!st.data) {
        SetError("refinement criterion field has no data in this MeshData");
        return PTH_FAIL;
      }
      TagArgs a;
      a.q = st.data;
      a.ncomp = fields[c.field].ncomp;
      a.comp = c.comp;
      for (int d = 0; d < 3; ++d) {
        a.e[d] = topo.fine.ent(d);
        a.lo[d] = topo.fine.lo(d);
        a.n[d] = topo.fine.n[d];
      }
      a.ndim = 1 + (a.n[1] > 1) + (a.n[2] > 1);
      a.method = c.method;
      ProfScope prof("refinement_tag", stream);
      hipLaunchKernelGGL(DerivativeMaxKernel, dim3(nb), dim3(kTagWG), 0, stream, a, amr.maxd, nc,
                         q);
      PTH_HIP(hipGetLastError());
    }
    PTH_HIP(hipMemcpyAsync(maxd.data(), amr.maxd, sizeof(double) * need, hipMemcpyDeviceToHost,
                           stream));
    PTH_HIP(hipStreamSynchronize(stream));
  }
  // CheckAllRefinement: the package default recommends derefinement; each criterion's refine
  // is withheld at or above its max_level (a logical level)
  const int root = topo.RootLevel();
  for (int b = 0; b < nb; ++b) {
    int delta = -1;
    for (int q = 0; q < nc; ++q) {
      const AmrCriterion &c = amr.crit[q];
      const double m = maxd[static_cast<size_t>(b) * nc + q];
      int t = m > c.refine_tol ? 1 : (m < c.derefine_tol ? -1 : 0);
      if (t == 1 && root + topo.blocks[b].at.level >= c.max_level) t = 0;
      delta = std::max(delta, t);
      if (delta == 1) break;
    }
    SetRefinement(b, delta);
  }
  return PTH_COMPLETE;
}

void Mesh::Adopt(Mesh &nm) {
  std::swap(topo, nm.topo);
  for (auto &kv : data) {
    MeshData *mine = kv.second.get();
    MeshData *fresh = nm.data.at(kv.first).get();
    std::swap(mine->store, fresh->store);
    for (int q = 0; q < 10; ++q) std::swap(mine->cache[q], fresh->cache[q]);
  }
  std::swap(one_copy, nm.one_copy);
  std::swap(arena, nm.arena);
  std::swap(arena_size, nm.arena_size);
  std::swap(arena_base, nm.arena_base);
  std::swap(arena_total, nm.arena_total);
  std::swap(channel_offset, nm.channel_offset);
  std::swap(geom, nm.geom);
  std::swap(coef, nm.coef);
  std::swap(face, nm.face);
  std::swap(face_V, nm.face_V);
  std::swap(graph, nm.graph);
  std::swap(graph_exec, nm.graph_exec);
  std::swap(scratch, nm.scratch);
  std::swap(scratch_n, nm.scratch_n);
  std::swap(pinned, nm.pinned);
  std::swap(flxcor, nm.flxcor);
  std::swap(same_level, nm.same_level);
  for (BufState &st : state) st = BufState::kStale;
  flxcor_state = BufState::kStale;
  ++generation;
}

pth_status Mesh::Adapt(int *modified) {
  *modified = 0;
  if (!finalized || !amr.on) {
    SetError("pth_mesh_adapt needs a finalized adaptive mesh");
    return PTH_FAIL;
  }
  const int nb = static_cast<int>(topo.blocks.size());
  // UpdateMeshBlockTree: blocks flagged 1 refine (gid order); derefinement candidates are the
  // parents whose nleaf daughters follow each other in the gid-ordered list of blocks flagged
  // -1, sorted finest first over all but the last entry (the reference's sort bound)
  const int nleaf = 1 << topo.ndim;
  std::vector<Lattice> lref, lderef;
  for (int b = 0; b < nb; ++b) {
    if (amr.flag[b] == 1) lref.push_back(topo.blocks[b].at);
    if (amr.flag[b] == -1) lderef.push_back(topo.blocks[b].at);
  }
  if (lref.empty() && static_cast<int>(lderef.size()) < nleaf) return PTH_COMPLETE;
  std::vector<Lattice> parents;
  if (static_cast<int>(lderef.size()) >= nleaf) {
    const int lj = topo.active[1] ? 1 : 0, lk = topo.active[2] ? 1 : 0;
    for (size_t n = 0; n < lderef.size(); ++n) {
      const Lattice &a = lderef[n];
      if ((a.x[0] & 1) || (a.x[1] & 1) || (a.x[2] & 1)) continue;
      size_t r = n;
      int hits = 0;
      for (int k = 0; k <= lk; ++k)
        for (int j = 0; j <= lj; ++j)
          for (int i = 0; i <= 1; ++i) {
            if (r >= lderef.size()) continue;
            const Lattice &c = lderef[r++];
            if (c.level == a.level && c.x[0] == a.x[0] + i && c.x[1] == a.x[1] + j &&
                c.x[2] == a.x[2] + k)
              hits++;
          }
      if (hits == nleaf) parents.push_back(Parent(a));
    }
  }
  if (parents.size() > 1)
    std::stable_sort(parents.begin(), parents.end() - 1,
                     [](const Lattice &x, const Lattice &y) { return x.level > y.level; });
  Topology t = topo;
  int nnew = 0, ndel = 0;
  for (const Lattice &a : lref) nnew += t.Refine(a);
  for (const Lattice &p : parents) ndel += t.Derefine(p);
  if (nnew == 0 && ndel == 0) return PTH_COMPLETE;
  t.blocks.clear();
  t.id.clear();
  const std::string err = t.Finalize();
  if (!err.empty()) {
    SetError(err);
    return PTH_FAIL;
  }
  if (!t.Nested()) {
    SetError("internal: remeshed tree is not properly nested");
    return PTH_FAIL;
  }
  // the new mesh beside the old one: same fields and settings, every MeshData allocated
  std::unique_ptr<Mesh> nm(new Mesh);
  nm->topo = std::move(t);
  nm->fields = fields;
  nm->stream = stream;
  nm->pack_mode = pack_mode;
  nm->planned = true;
  if (nm->Finalize() != PTH_COMPLETE) return PTH_FAIL;
  for (auto &kv : data)
    if (!nm->Get(kv.first)) return PTH_FAIL;
  const Topology &nt = nm->topo;
  const int nbn = static_cast<int>(nt.blocks.size());
  std::vector<Origin> orig(nbn);
  for (int n = 0; n < nbn; ++n) {
    const Lattice &at = nt.blocks[n].at;
    Origin &o = orig[n];
    auto it = topo.id.find(at);
    if (it != topo.id.end()) {
      o.kind = 0;
      o.old[0] = it->second;
      continue;
    }
    if (at.level > 0) {
      auto ip = topo.id.find(Parent(at));
      if (ip != topo.id.end()) {
        o.kind = 1;
        o.old[0] = ip->second;
        for (int d = 0; d < 3; ++d) o.ox[0][d] = static_cast<int>(at.x[d] & 1);
        continue;
      }
    }
    o.kind = 2;
    o.n = 0;
    for (int k = 0; k < (nt.active[2] ? 2 : 1); ++k)
      for (int j = 0; j < (nt.active[1] ? 2 : 1); ++j)
        for (int i = 0; i < 2; ++i) {
          Lattice c;
          c.level = at.level + 1;
          c.x[0] = 2 * at.x[0] + i;
          c.x[1] = nt.active[1] ? 2 * at.x[1] + j : 0;
          c.x[2] = nt.active[2] ? 2 * at.x[2] + k : 0;
          auto ic = topo.id.find(c);
          if (ic == topo.id.end() || o.n >= 8) {
            SetError("internal: a block changed by more than one level in one remesh");
            return PTH_FAIL;
          }
          o.old[o.n] = ic->second;
          o.ox[o.n][0] = i;
          o.ox[o.n][1] = c.x[1] & 1;
          o.ox[o.n][2] = c.x[2] & 1;
          o.n++;
        }
  }
  // region tables: restriction on the old list, box copies old -> new, prolongation on the new
  Table rt, ct, pt, pl;  // pl: prolongation of fields registered with the linear operator
  const Extent &F = topo.fine, &Cx = topo.coarse;
  const std::int64_t fcells = F.cells(), ccells = Cx.cells();
  auto add_field = [&](const FieldStore &os, const FieldStore &ns, int nc, bool linear) -> bool {
    if (!os.data || !ns.data) return true;
    if (!os.coarse || !ns.coarse) return false;
    for (int n = 0; n < nbn; ++n) {
      const Origin &o = orig[n];
      double *dst = ns.data + static_cast<std::int64_t>(n) * nc * fcells;
      if (o.kind == 0) {
        const int lo[3] = {0, 0, 0}, sz[3] = {F.ent(0), F.ent(1), F.ent(2)};
        ct.host.push_back(MakeBox(os.data + static_cast<std::int64_t>(o.old[0]) * nc * fcells,
                                  dst, lo, sz, nc, n, F, F));
      } else if (o.kind == 1) {
        int lo[3], sz[3], plo[3], psz[3];
        for (int d = 0; d < 3; ++d) {
          lo[d] = o.ox[0][d] ? F.n[d] / 2 : 0;
          sz[d] = Cx.ent(d);
          const int h = nt.active[d] ? nt.ng / 2 : 0;
          plo[d] = Cx.lo(d) - h;
          psz[d] = Cx.n[d] + 2 * h;
        }
        double *cdst = ns.coarse + static_cast<std::int64_t>(n) * nc * ccells;
        ct.host.push_back(MakeBox(os.data + static_cast<std::int64_t>(o.old[0]) * nc * fcells,
                                  cdst, lo, sz, nc, n, F, Cx));
        (linear ? pl : pt).host.push_back(MakeBox(dst, cdst, plo, psz, nc, n, F, Cx));
      } else {
        for (int q = 0; q < o.n; ++q) {
          double *fsrc = os.data + static_cast<std::int64_t>(o.old[q]) * nc * fcells;
          double *csrc = os.coarse + static_cast<std::int64_t>(o.old[q]) * nc * ccells;
          const int clo[3] = {Cx.lo(0), Cx.lo(1), Cx.lo(2)};
          const int cn[3] = {Cx.n[0], Cx.n[1], Cx.n[2]};
          rt.host.push_back(MakeBox(fsrc, csrc, clo, cn, nc, o.old[q], F, Cx));
          std::int64_t off = 0;
          const std::int64_t stride[3] = {1, F.ent(0), static_cast<std::int64_t>(F.ent(0)) * F.ent(1)};
          for (int d = 0; d < 3; ++d) off += (F.lo(d) + o.ox[q][d] * Cx.n[d]) * stride[d];
          ct.host.push_back(MakeBox(csrc, dst + off, clo, cn, nc, n, Cx, F));
        }
      }
    }
    return true;
  };
  for (auto &kv : data) {
    MeshData &om = *kv.second;
    MeshData &nmd = *nm->data.at(kv.first);
    for (size_t f = 0; f < fields.size(); ++f) {
      if (fields[f].flags & PTH_ONE_COPY) continue;
      if (!add_field(om.store[f], nmd.store[f], fields[f].ncomp,
                     fields[f].prolong == PTH_PROLONG_LINEAR)) {
        SetError("internal: remesh needs coarse buffers");
        return PTH_FAIL;
      }
    }
  }
  for (size_t f = 0; f < fields.size(); ++f)
    if (fields[f].flags & PTH_ONE_COPY)
      if (!add_field(one_copy[f], nm->one_copy[f], fields[f].ncomp,
                     fields[f].prolong == PTH_PROLONG_LINEAR)) {
        SetError("internal: remesh needs coarse buffers");
        return PTH_FAIL;
      }
  auto fail = [&]() {
    rt.Free();
    ct.Free();
    pt.Free();
    pl.Free();
    return PTH_FAIL;
  };
  rt.item_elems = pt.item_elems = pl.item_elems = kProResItemElems;
  if (rt.Upload() != PTH_COMPLETE || ct.Upload() != PTH_COMPLETE || pt.Upload() != PTH_COMPLETE ||
      pl.Upload() != PTH_COMPLETE)
    return fail();
  if (LaunchRestrictT(rt, topo, geom, stream) != PTH_COMPLETE) return fail();
  if (LaunchCopy(ct, stream) != PTH_COMPLETE) return fail();
  if (LaunchProlongT(pt, nt, nm->geom, stream, true) != PTH_COMPLETE) return fail();
  if (LaunchProlongT(pl, nt, nm->geom, stream, false) != PTH_COMPLETE) return fail();
  if (hipStreamSynchronize(stream) != hipSuccess) {
    SetError("remesh transfer failed on the device");
    return fail();
  }
  rt.Free();
  ct.Free();
  pt.Free();
  pl.Free();
  // flags: moved blocks keep theirs, new blocks start at 0 (MeshRefinement is rebuilt)
  std::vector<int> fl(nbn, 0), cnt(nbn, 0);
  for (int n = 0; n < nbn; ++n)
    if (orig[n].kind == 0) {
      fl[n] = amr.flag[orig[n].old[0]];
      cnt[n] = amr.count[orig[n].old[0]];
    }
  Adopt(*nm);
  amr.flag.swap(fl);
  amr.count.swap(cnt);
  *modified = 1;
  return PTH_COMPLETE;
}

}  // namespace pth

using namespace pth;

extern "C" {

pth_status pth_mesh_set_adaptive(pth_mesh *mh, int numlevel, int derefine_count) {
  Mesh *m = Impl(mh);
  if (m->planned) {
    SetError("pth_mesh_set_adaptive must precede pth_mesh_plan / pth_mesh_finalize");
    return PTH_FAIL;
  }
  if (numlevel < 1 || derefine_count < 1) {
    SetError("numlevel and derefine_count must be >= 1");
    return PTH_FAIL;
  }
  m->amr.on = true;
  m->amr.numlevel = numlevel;
  m->amr.deref_threshold = derefine_count;
  m->topo.multilevel = true;
  return PTH_COMPLETE;
}

pth_status pth_mesh_add_refinement_criterion(pth_mesh *mh, int method, const char *field,
                                             int comp, double refine_tol, double derefine_tol,
                                             int max_level) {
  Mesh *m = Impl(mh);
  if (!m->amr.on) {
    SetError("refinement criteria need an adaptive mesh (pth_mesh_set_adaptive)");
    return PTH_FAIL;
  }
  const int f = m->FieldIndex(field ? field : "");
  if (f < 0 || comp < 0 || comp >= m->fields[f].ncomp) {
    SetError("refinement criterion: unknown field or component");
    return PTH_FAIL;
  }
  if (method != 1 && method != 2) {
    SetError("refinement method must be 1 (derivative_order_1) or 2 (derivative_order_2)");
    return PTH_FAIL;
  }
  AmrCriterion c;
  c.method = method;
  c.field = f;
  c.comp = comp;
  c.refine_tol = refine_tol;
  c.derefine_tol = derefine_tol;
  c.max_level = (max_level <= 0 || max_level > m->amr.numlevel) ? m->amr.numlevel : max_level;
  m->amr.crit.push_back(c);
  return PTH_COMPLETE;
}

static bool AmrReady(Mesh *m) {
  if (!m->finalized || !m->amr.on) {
    SetError("needs a finalized adaptive mesh");
    return false;
  }
  const size_t nb = m->topo.blocks.size();
  if (m->amr.flag.size() != nb) {
    m->amr.flag.assign(nb, 0);
    m->amr.count.assign(nb, 0);
  }
  return true;
}

pth_status pth_refinement_tag(pth_meshdata *mdh) {
  MeshData *md = Impl(mdh);
  if (!AmrReady(md->mesh)) return PTH_FAIL;
  return md->mesh->Tag(*md);
}

pth_status pth_mesh_set_refinement(pth_mesh *mh, int gid, int tag) {
  Mesh *m = Impl(mh);
  if (!AmrReady(m)) return PTH_FAIL;
  if (gid < 0 || gid >= static_cast<int>(m->topo.blocks.size())) {
    SetError("block id out of range");
    return PTH_FAIL;
  }
  m->SetRefinement(gid, tag);
  return PTH_COMPLETE;
}

pth_status pth_mesh_refinement_flags(const pth_mesh *mh, int *flags, int *counts) {
  const Mesh *m = Impl(mh);
  if (!m->finalized || !m->amr.on) {
    SetError("needs a finalized adaptive mesh");
    return PTH_FAIL;
  }
  // blocks not tagged yet (flags not sized) report 0, as a new MeshRefinement holds
  const size_t nb = m->topo.blocks.size();
  const bool sized = m->amr.flag.size() == nb;
  for (size_t b = 0; b < nb; ++b) {
    if (flags) flags[b] = sized ? m->amr.flag[b] : 0;
    if (counts) counts[b] = sized ? m->amr.count[b] : 0;
  }
  return PTH_COMPLETE;
}

int pth_mesh_root_level(const pth_mesh *mh) { return Impl(mh)->topo.RootLevel(); }

pth_status pth_mesh_adapt(pth_mesh *mh, int *modified) {
  Mesh *m = Impl(mh);
  int mod = 0;
  if (modified) *modified = 0;
  if (!AmrReady(m)) return PTH_FAIL;
  if (m->Adapt(&mod) != PTH_COMPLETE) return PTH_FAIL;
  if (mod) {
    // CommunicateBoundaries on "base": ghosts of the new blocks from their neighbours
    auto it = m->data.find("base");
    if (it != m->data.end() &&
        pth_exchange_fused(reinterpret_cast<pth_meshdata *>(it->second.get()), PTH_BT_ANY) !=
            PTH_COMPLETE)
      return PTH_FAIL;
    PTH_HIP(hipStreamSynchronize(m->stream));
  }
  if (modified) *modified = mod;
  return PTH_COMPLETE;
}

}  // extern "C"
