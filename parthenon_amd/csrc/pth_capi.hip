// C ABI of parthenon_amd: mesh and MeshData management, boundary communication tasks and
// the Burgers package tasks. Host-side decisions (table builds, buffer-state flips) happen
// here; every data operation is a device kernel on the mesh's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "pth_burgers.hpp"

namespace pth {

namespace {
thread_local std::string g_error;
}

void SetError(const std::string &msg) { g_error = msg; }
const char *ErrorText() { return g_error.c_str(); }

Profiler &Prof() {
  static Profiler p;
  return p;
}

hipEvent_t Profiler::Take() {
  if (pool.empty()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  hipEvent_t e = pool.back();
  pool.pop_back();
  return e;
}

void Profiler::Clear() {
  for (auto &r : records) {
    pool.push_back(r.start);
    pool.push_back(r.stop);
  }
  records.clear();
}

ProfScope::ProfScope(const char *name, hipStream_t s) : name_(name), s_(s) {
  if (!Prof().on) return;
  a_ = Prof().Take();
  if (a_) (void)hipEventRecord(a_, s_);
}

ProfScope::~ProfScope() {
  if (!a_) return;
  hipEvent_t b = Prof().Take();
  if (!b) return;
  (void)hipEventRecord(b, s_);
  Prof().records.push_back(ProfRecord{name_, a_, b});
}

MeshData::~MeshData() {
  for (auto &s : store) {
    if (!s.owned) continue;
    if (s.data) (void)hipFree(s.data);
    if (s.coarse) (void)hipFree(s.coarse);
    if (s.flux) (void)hipFree(s.flux);
  }
  for (auto &c : cache) {
    c.pack.Free();
    c.unpack.Free();
    c.restrict_send.Free();
    c.restrict_set.Free();
    c.prolong.Free();
    c.prolong_linear.Free();
    c.bc[0].Free();
    c.bc[1].Free();
    c.pack_rest.Free();
    c.unpack_rest.Free();
    c.pack_cross.Free();
    c.unpack_cross.Free();
    if (c.shell_order) (void)hipFree(c.shell_order);
    c.shell_order = nullptr;
  }
  for (auto &kv : packs) kv.second.Free();
}

void PackEntry::Free() {
  if (dev) (void)hipFree(dev);
  dev = nullptr;
}

Mesh::~Mesh() {
  data.clear();
  for (auto &kv : flxcor) kv.second.Free();
  for (auto &s : one_copy) {
    if (s.data) (void)hipFree(s.data);
    if (s.coarse) (void)hipFree(s.coarse);
    if (s.flux) (void)hipFree(s.flux);
  }
  if (arena_base) (void)hipFree(arena_base);
  if (same_level) (void)hipFree(same_level);
  if (split_list) (void)hipFree(split_list);
  if (geom) (void)hipFree(geom);
  if (coef) (void)hipFree(coef);
  if (face) (void)hipFree(face);
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (run_stream) (void)hipStreamDestroy(run_stream);
  for (hipEvent_t e : order_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : fork_ev)
    if (e) (void)hipEventDestroy(e);
  if (side_stream) (void)hipStreamDestroy(side_stream);
  if (dtvar) (void)hipFree(dtvar);
  if (scratch) (void)hipFree(scratch);
  if (pinned) (void)hipHostFree(pinned);
  if (amr.maxd) (void)hipFree(amr.maxd);
}

int Mesh::FieldIndex(const std::string &n) const {
  for (size_t q = 0; q < fields.size(); ++q)
    if (fields[q].name == n) return static_cast<int>(q);
  return -1;
}

}  // namespace pth

using namespace pth;

namespace {

pth_status Alloc(double **p, std::int64_t n) {
  *p = nullptr;
  if (n <= 0) return PTH_COMPLETE;
  PTH_HIP(hipMalloc(p, sizeof(double) * n));
  PTH_HIP(hipMemset(*p, 0, sizeof(double) * n));
  return PTH_COMPLETE;
}

pth_status AllocStore(const Mesh &m, const FieldSpec &f, FieldStore &s) {
  const std::int64_t nb = static_cast<std::int64_t>(m.topo.blocks.size());
  s.ndata = nb * f.ncomp * m.topo.fine.cells();
  s.ncoarse = (m.topo.multilevel || m.force_coarse) ? nb * f.ncomp * m.topo.coarse.cells() : 0;
  s.nflux = (f.flags & PTH_WITH_FLUXES) ? 3 * s.ndata : 0;
  if (Alloc(&s.data, s.ndata) != PTH_COMPLETE) return PTH_FAIL;
  if (Alloc(&s.coarse, s.ncoarse) != PTH_COMPLETE) return PTH_FAIL;
  if (Alloc(&s.flux, s.nflux) != PTH_COMPLETE) return PTH_FAIL;
  return PTH_COMPLETE;
}

Region MakeRegion(double *a, double *b, const Box &box, int nc, int gid, const Extent &ea,
                  const Extent *eb) {
  Region r;
  std::memset(&r, 0, sizeof(r));
  r.a = a;
  r.b = b;
  r.nc = nc;
  r.gid = gid;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = box.lo[d];
    r.n[d] = box.hi[d] - box.lo[d] + 1;
    r.magic[d] = Magic(r.n[d]);
  }
  r.astride[0] = ea.ent(0);
  r.astride[1] = ea.ent(0) * ea.ent(1);
  r.astride[2] = static_cast<int>(ea.cells());
  if (eb) {
    r.bstride[0] = eb->ent(0);
    r.bstride[1] = eb->ent(0) * eb->ent(1);
    r.bstride[2] = static_cast<int>(eb->cells());
  }
  return r;
}

bool IsExchangeType(int bt) { return bt == PTH_BT_ANY || bt == PTH_BT_LOCAL || bt == PTH_BT_NONLOCAL; }

}  // namespace

namespace pth {

pth_status Mesh::Plan() {
  if (planned) return PTH_COMPLETE;
  const std::string err = topo.Finalize();
  if (!err.empty()) {
    SetError(err);
    return PTH_FAIL;
  }
  planned = true;
  return PTH_COMPLETE;
}

pth_status Mesh::Finalize() {
  if (finalized) {
    SetError("mesh already finalized");
    return PTH_FAIL;
  }
  if (Plan() != PTH_COMPLETE) return PTH_FAIL;
  const int nb = static_cast<int>(topo.blocks.size());
  // per-block geometry for prolongation and metrics for the stencils
  std::vector<DeviceBlockGeom> g(nb);
  std::vector<BlockCoefs> c(nb);
  for (int b = 0; b < nb; ++b) {
    const Block &B = topo.blocks[b];
    for (int d = 0; d < 3; ++d) {
      g[b].x0[d] = B.x0[d];
      g[b].dx[d] = B.dx[d];
      g[b].cx0[d] = B.cx0[d];
      g[b].cdx[d] = B.cdx[d];
      c[b].dx[d] = B.dx[d];
      c[b].area[d] = B.area[d];
    }
    c[b].vol = B.vol;
    c[b].pad = 0.0;
  }
  PTH_HIP(hipMalloc(&geom, sizeof(DeviceBlockGeom) * nb));
  PTH_HIP(hipMemcpy(geom, g.data(), sizeof(DeviceBlockGeom) * nb, hipMemcpyHostToDevice));
  PTH_HIP(hipMalloc(&coef, sizeof(BlockCoefs) * nb));
  PTH_HIP(hipMemcpy(coef, c.data(), sizeof(BlockCoefs) * nb, hipMemcpyHostToDevice));
  // OneCopy fields: one storage shared by every MeshData
  one_copy.assign(fields.size(), FieldStore());
  for (size_t f = 0; f < fields.size(); ++f)
    if (fields[f].flags & PTH_ONE_COPY)
      if (AllocStore(*this, fields[f], one_copy[f]) != PTH_COMPLETE) return PTH_FAIL;
  // communication channels: one arena for every FillGhost field, one contiguous slice per
  // (sender, receiver, field, location), laid out receiver-block-major: the unpack then writes
  // each receiving brick's ghost shell in one pass (partial cache lines merge in L2), and the
  // pack gathers from the senders
  arena.assign(fields.size(), nullptr);
  arena_size.assign(fields.size(), 0);
  channel_offset.clear();
  std::map<Channel, std::int64_t> slice;
  for (int b = 0; b < nb; ++b) {
    const Block &B = topo.blocks[b];
    for (size_t f = 0; f < fields.size(); ++f) {
      if (!(fields[f].flags & PTH_FILL_GHOST)) continue;
      for (const Link &l : B.links) {
        const Side side = l.is_cell() ? Side::kSendInterior : Side::kSendBoundary;
        slice[Channel(b, l.gid, static_cast<int>(f), l.loc_index())] =
            topo.Indices(B, l, side, false).count() * fields[f].ncomp;
      }
    }
  }
  // every receiving region must match the size of the slice its sender packs
  std::int64_t off = 0;
  for (int b = 0; b < nb; ++b) {
    const Block &B = topo.blocks[b];
    for (size_t f = 0; f < fields.size(); ++f) {
      if (!(fields[f].flags & PTH_FILL_GHOST)) continue;
      for (const Link &l : B.links) {
        const Side side = l.is_cell() ? Side::kRecvInterior : Side::kRecvBoundary;
        const Box box = topo.Indices(B, l, side, false);
        const Channel key(l.gid, b, static_cast<int>(f), l.reverse_index());
        auto it = slice.find(key);
        if (it == slice.end() || it->second != box.count() * fields[f].ncomp ||
            channel_offset.count(key)) {
          SetError("receiving region does not match the slice its sender packs");
          return PTH_FAIL;
        }
        channel_offset[key] = off;
        off += it->second;
      }
    }
  }
  if (channel_offset.size() != slice.size()) {
    SetError("a sending region has no receiver");
    return PTH_FAIL;
  }
  double *all = nullptr;
  if (Alloc(&all, off) != PTH_COMPLETE) return PTH_FAIL;
  arena_base = all;
  arena_total = off;
  for (size_t f = 0; f < fields.size(); ++f) {
    arena[f] = all;
    arena_size[f] = off;
  }
  scratch_n = std::max<std::int64_t>(8192, FusedUpdatePartials(topo));
  PTH_HIP(hipMalloc(&scratch, sizeof(double) * (scratch_n + 8)));
  PTH_HIP(hipHostMalloc(&pinned, sizeof(double) * 8, hipHostMallocDefault));
  finalized = true;
  return PTH_COMPLETE;
}

MeshData *Mesh::Get(const std::string &name) {
  auto it = data.find(name);
  if (it != data.end()) return it->second.get();
  std::unique_ptr<MeshData> md(new MeshData);
  md->mesh = this;
  md->name = name;
  md->store.assign(fields.size(), FieldStore());
  for (size_t f = 0; f < fields.size(); ++f) {
    if (fields[f].flags & PTH_ONE_COPY) {
      md->store[f] = one_copy[f];
      md->store[f].owned = false;
    } else if (AllocStore(*this, fields[f], md->store[f]) != PTH_COMPLETE) {
      return nullptr;
    }
  }
  MeshData *raw = md.get();
  data[name] = std::move(md);
  return raw;
}

// Builds the pack / unpack / restriction / prolongation region tables of md for bound type bt.
pth_status Mesh::BuildCache(MeshData &md, int bt) {
  BoundCache &c = md.cache[bt];
  if (c.built) return PTH_COMPLETE;
  c.pack.host.clear();
  c.unpack.host.clear();
  c.restrict_send.host.clear();
  c.restrict_set.host.clear();
  c.prolong.host.clear();
  c.prolong_linear.host.clear();
  c.nbuffers = 0;
  if (bt == PTH_BT_NONLOCAL) {  // one rank: no neighbour is remote
    c.built = true;
    return PTH_COMPLETE;
  }
  if (!IsExchangeType(bt)) {
    SetError("bound type not available for cell-variable exchange in this build");
    return PTH_FAIL;
  }
  const Extent &fe = topo.fine, &ce = topo.coarse;
  std::vector<char> on_coarse;  // pack/unpack region q addresses the coarse buffer
  for (int b = 0; b < static_cast<int>(topo.blocks.size()); ++b) {
    const Block &B = topo.blocks[b];
    for (size_t f = 0; f < fields.size(); ++f) {
      if (!(fields[f].flags & PTH_FILL_GHOST)) continue;
      const int nc = fields[f].ncomp;
      FieldStore &S = md.store[f];
      double *fine = S.data + static_cast<std::int64_t>(b) * nc * fe.cells();
      double *crs = S.coarse ? S.coarse + static_cast<std::int64_t>(b) * nc * ce.cells() : nullptr;
      for (const Link &l : B.links) {
        const bool coarser = l.level < B.at.level;
        const Extent &ve = coarser ? ce : fe;
        double *var = coarser ? crs : fine;
        // send region
        const Side ss = l.is_cell() ? Side::kSendInterior : Side::kSendBoundary;
        const Box sb = topo.Indices(B, l, ss, false);
        const std::int64_t so =
            channel_offset.at(Channel(b, l.gid, static_cast<int>(f), l.loc_index()));
        c.pack.host.push_back(MakeRegion(var, arena[f] + so, sb, nc, b, ve, nullptr));
        // receive region
        const Side rs = l.is_cell() ? Side::kRecvInterior : Side::kRecvBoundary;
        const Box rb = topo.Indices(B, l, rs, false);
        const std::int64_t ro =
            channel_offset.at(Channel(l.gid, b, static_cast<int>(f), l.reverse_index()));
        c.unpack.host.push_back(MakeRegion(var, arena[f] + ro, rb, nc, b, ve, nullptr));
        on_coarse.push_back(coarser ? 1 : 0);
        ++c.nbuffers;
        if (!topo.multilevel) continue;
        if (coarser) {
          c.restrict_send.host.push_back(MakeRegion(
              fine, crs, topo.Indices(B, l, Side::kSendBoundary, true), nc, b, fe, &ce));
          (fields[f].prolong == PTH_PROLONG_LINEAR ? c.prolong_linear : c.prolong)
              .host.push_back(MakeRegion(fine, crs,
                                         topo.Indices(B, l, Side::kRecvBoundary, true), nc, b,
                                         fe, &ce));
        } else if (B.coarser_neighbor && B.at.level > 0) {
          c.restrict_set.host.push_back(MakeRegion(
              fine, crs, topo.Indices(B, l, Side::kRecvBoundary, true), nc, b, fe, &ce));
        }
      }
    }
  }
  // Restriction before the send: the regions of one block's links toward coarser neighbours
  // overlap (a face region contains the halves of its edge and corner regions), and every
  // coarse cell restricts to the same value whichever region asks for it; restrict each
  // block's union once, as disjoint boxes.
  // Restriction after the set: the reference restricts every same-level and finer ghost
  // region of a block with a coarser neighbour, "a superset of the boundaries that actually
  // need to be restricted" (src/bvals/comms/bnd_info.cpp:409-411). The prolongation reads
  // the coarse buffer only within one coarse cell of its own boxes (the minmod stencil), so
  // on meshes whose active dimensions are all periodic the set is cut to those cells. With a
  // physical boundary the coarse-buffer conditions can copy other halo cells into that
  // reach, and the whole superset is kept.
  bool all_periodic = true;
  for (int d = 0; d < 3; ++d) all_periodic = all_periodic && (topo.periodic[d] || !topo.active[d]);
  auto merge = [&](std::vector<Region> &host, const std::vector<Region> *reach) {
    std::vector<Region> merged;
    std::map<std::pair<double *, double *>, std::vector<const Region *>> by_block;
    std::map<std::pair<double *, double *>, std::vector<const Region *>> reach_of;
    std::vector<std::pair<double *, double *>> order;
    for (const Region &r : host) {
      auto key = std::make_pair(r.a, r.b);
      if (!by_block.count(key)) order.push_back(key);
      by_block[key].push_back(&r);
    }
    if (reach)
      for (const Region &r : *reach) reach_of[std::make_pair(r.a, r.b)].push_back(&r);
    const int E0 = ce.ent(0), E1 = ce.ent(1), E2 = ce.ent(2);
    const int E[3] = {E0, E1, E2};
    std::vector<char> mask(static_cast<size_t>(E0) * E1 * E2);
    auto at = [&](int i, int j, int k) -> char & {
      return mask[(static_cast<size_t>(k) * E1 + j) * E0 + i];
    };
    for (const auto &key : order) {
      const auto &rs = by_block[key];
      std::fill(mask.begin(), mask.end(), 0);
      // 3 = wanted; with a reach, only the cells it covers become 1 (to restrict)
      const char want = reach ? 3 : 1;
      for (const Region *r : rs)
        for (int k = r->lo[2]; k < r->lo[2] + r->n[2]; ++k)
          for (int j = r->lo[1]; j < r->lo[1] + r->n[1]; ++j)
            for (int i = r->lo[0]; i < r->lo[0] + r->n[0]; ++i) at(i, j, k) = want;
      if (reach) {
        auto it = reach_of.find(key);
        if (it != reach_of.end())
          for (const Region *p : it->second) {
            int lo[3], hi[3];
            for (int d = 0; d < 3; ++d) {
              lo[d] = std::max(p->lo[d] - 1, 0);
              hi[d] = std::min(p->lo[d] + p->n[d], E[d] - 1);
            }
            for (int k = lo[2]; k <= hi[2]; ++k)
              for (int j = lo[1]; j <= hi[1]; ++j)
                for (int i = lo[0]; i <= hi[0]; ++i)
                  if (at(i, j, k) == 3) at(i, j, k) = 1;
          }
      }
      const Region &r0 = *rs.front();
      for (int k = 0; k < E2; ++k)
        for (int j = 0; j < E1; ++j)
          for (int i = 0; i < E0; ++i) {
            if (at(i, j, k) != 1) continue;
            int i1 = i;
            while (i1 + 1 < E0 && at(i1 + 1, j, k) == 1) ++i1;
            auto row_ok = [&](int jj, int kk) {
              for (int ii = i; ii <= i1; ++ii)
                if (at(ii, jj, kk) != 1) return false;
              return true;
            };
            int j1 = j;
            while (j1 + 1 < E1 && row_ok(j1 + 1, k)) ++j1;
            int k1 = k;
            for (bool grow = true; grow && k1 + 1 < E2;) {
              for (int jj = j; jj <= j1 && grow; ++jj) grow = row_ok(jj, k1 + 1);
              if (grow) ++k1;
            }
            for (int kk = k; kk <= k1; ++kk)
              for (int jj = j; jj <= j1; ++jj)
                for (int ii = i; ii <= i1; ++ii) at(ii, jj, kk) = 2;
            Box bx;
            bx.lo[0] = i;
            bx.lo[1] = j;
            bx.lo[2] = k;
            bx.hi[0] = i1;
            bx.hi[1] = j1;
            bx.hi[2] = k1;
            merged.push_back(MakeRegion(r0.a, r0.b, bx, r0.nc, r0.gid, fe, &ce));
          }
    }
    host.swap(merged);
  };
  merge(c.restrict_send.host, nullptr);
  if (all_periodic) {
    std::vector<Region> reach = c.prolong.host;
    reach.insert(reach.end(), c.prolong_linear.host.begin(), c.prolong_linear.host.end());
    merge(c.restrict_set.host, &reach);
  }
  // Every element a kernel will touch must lie inside its array: pack/unpack boxes inside the
  // field brick, restriction children and prolongation stencils inside fine and coarse bricks.
  auto inside = [](const Region &r, const Extent &e, int halo_lo, int halo_hi, int scale,
                   int base_lo, int base_fine) {
    for (int d = 0; d < 3; ++d) {
      if (e.ent(d) == 1) {
        if (r.lo[d] != 0 || r.n[d] != 1) return false;
        continue;
      }
      const int lo = scale == 1 ? r.lo[d] - halo_lo : (r.lo[d] - base_lo) * 2 + base_fine;
      const int hi = scale == 1 ? r.lo[d] + r.n[d] - 1 + halo_hi
                                : (r.lo[d] + r.n[d] - 1 - base_lo) * 2 + base_fine + 1;
      if (lo < 0 || hi >= e.ent(d)) return false;
    }
    return true;
  };
  bool ok = true;
  for (size_t q = 0; q < c.pack.host.size(); ++q) {
    const Extent &e = on_coarse[q] ? ce : fe;
    ok = ok && inside(c.pack.host[q], e, 0, 0, 1, 0, 0) && inside(c.unpack.host[q], e, 0, 0, 1, 0, 0);
  }
  for (const Table *t : {&c.restrict_send, &c.restrict_set})
    for (const Region &r : t->host)
      ok = ok && inside(r, ce, 0, 0, 1, 0, 0) && inside(r, fe, 0, 0, 2, ce.lo(0), fe.lo(0));
  for (const Table *t : {&c.prolong, &c.prolong_linear})
    for (const Region &r : t->host)
      ok = ok && inside(r, ce, 1, 1, 1, 0, 0) && inside(r, fe, 0, 0, 2, ce.lo(0), fe.lo(0));
  if (!ok) {
    SetError("internal: a boundary region exceeds its array extents");
    return PTH_FAIL;
  }
  // The order of regions within a pack or an unpack is free (their writes are disjoint), so
  // order both like the arena; then region q of the pack and of the unpack cover the same
  // arena slice, which the direct exchange walks pair by pair.
  for (Table *t : {&c.pack, &c.unpack})
    std::sort(t->host.begin(), t->host.end(),
              [](const Region &x, const Region &y) { return x.b < y.b; });
  for (Table *t : {&c.restrict_send, &c.restrict_set, &c.prolong, &c.prolong_linear})
    t->item_elems = kProResItemElems;
  for (Table *t : {&c.pack, &c.unpack, &c.restrict_send, &c.restrict_set, &c.prolong,
                   &c.prolong_linear})
    if (t->Upload() != PTH_COMPLETE) return PTH_FAIL;
  if (BuildBoundaryConditions(md, c) != PTH_COMPLETE) return PTH_FAIL;
  c.built = true;
  return PTH_COMPLETE;
}

pth_status Mesh::SameLevelNeighbors() {
  if (same_level) return PTH_COMPLETE;
  const int nb = static_cast<int>(topo.blocks.size());
  std::vector<int> t(static_cast<size_t>(27) * nb, -1);
  for (int b = 0; b < nb; ++b)
    for (const Link &l : topo.blocks[b].links)
      if (!l.is_cell() && l.level == topo.blocks[b].at.level) t[27 * b + l.loc_index()] = l.gid;
  PTH_HIP(hipMalloc(&same_level, sizeof(int) * t.size()));
  PTH_HIP(hipMemcpy(same_level, t.data(), sizeof(int) * t.size(), hipMemcpyHostToDevice));
  return PTH_COMPLETE;
}

// The direct exchange of every channel except the same-level channels of field f: the pack
// and unpack regions, both in arena order, pair up channel by channel (a channel's send and
// receive regions cover the same arena slice in the same element order); a pair whose two
// regions both address field f's data (not a coarse buffer) is a same-level channel of f.
pth_status Mesh::BuildRestExchange(MeshData &md, int f, bool *ok) {
  *ok = false;
  if (BuildCache(md, PTH_BT_ANY) != PTH_COMPLETE) return PTH_FAIL;
  BoundCache &c = md.cache[PTH_BT_ANY];
  if (!CanExchangeDirect(c.pack, c.unpack) || c.pack.host.size() != c.unpack.host.size())
    return PTH_COMPLETE;
  if (c.rest_field == f) {
    *ok = true;
    return PTH_COMPLETE;
  }
  c.pack_rest.Free();
  c.unpack_rest.Free();
  const double *f0 = md.store[f].data, *f1 = f0 + md.store[f].ndata;
  std::vector<Region> sregs, rregs;
  for (size_t q = 0; q < c.pack.host.size(); ++q) {
    const Region &sr = c.pack.host[q], &rr = c.unpack.host[q];
    if (sr.b != rr.b) {
      SetError("internal: send and receive regions do not pair up");
      return PTH_FAIL;
    }
    const bool same = sr.a >= f0 && sr.a < f1 && rr.a >= f0 && rr.a < f1;
    if (same) continue;
    sregs.push_back(sr);
    rregs.push_back(rr);
  }
  for (Table *t : {&c.pack_rest, &c.unpack_rest}) {
    t->host = t == &c.pack_rest ? sregs : rregs;
    if (t->Upload() != PTH_COMPLETE) return PTH_FAIL;
  }
  if (c.pack_rest.total != c.unpack_rest.total) {
    SetError("internal: send and receive element counts differ");
    return PTH_FAIL;
  }
  c.rest_field = f;
  *ok = true;
  return PTH_COMPLETE;
}

// The direct exchange split for the ghost-shell kernel: the same-level channels of every
// field go to ShellKernel (one entry per field component), the other region pairs (the
// channels between levels, through coarse buffers) stay in pack_cross / unpack_cross. ok is
// false where the shell kernel does not apply (block shape) or the tables have no direct form.
pth_status Mesh::BuildCrossExchange(MeshData &md, bool *ok) {
  *ok = false;
  if (!ShellSupported(topo)) return PTH_COMPLETE;
  if (BuildCache(md, PTH_BT_ANY) != PTH_COMPLETE) return PTH_FAIL;
  BoundCache &c = md.cache[PTH_BT_ANY];
  if (!CanExchangeDirect(c.pack, c.unpack) || c.pack.host.size() != c.unpack.host.size())
    return PTH_COMPLETE;
  if (c.cross_built) {
    *ok = true;
    return PTH_COMPLETE;
  }
  if (SameLevelNeighbors() != PTH_COMPLETE) return PTH_FAIL;
  c.pack_cross.Free();
  c.unpack_cross.Free();
  c.shell.clear();
  if (c.shell_order) (void)hipFree(c.shell_order);
  c.shell_order = nullptr;
  {  // blocks in Morton order of their positions on the finest level's lattice
    const int nb = static_cast<int>(topo.blocks.size());
    int top = 0;
    for (const Block &B : topo.blocks) top = std::max(top, B.at.level);
    auto spread = [](std::uint64_t v) {  // bits of v to every third bit
      std::uint64_t r = 0;
      for (int bit = 0; bit < 21; ++bit) r |= ((v >> bit) & 1ull) << (3 * bit);
      return r;
    };
    std::vector<std::pair<std::uint64_t, int>> key(nb);
    for (int b = 0; b < nb; ++b) {
      const Lattice &L = topo.blocks[b].at;
      const int sh = top - L.level;
      key[b] = {spread(static_cast<std::uint64_t>(L.x[0]) << sh) |
                    (spread(static_cast<std::uint64_t>(L.x[1]) << sh) << 1) |
                    (spread(static_cast<std::uint64_t>(L.x[2]) << sh) << 2),
                b};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> order(nb);
    for (int b = 0; b < nb; ++b) order[b] = key[b].second;
    PTH_HIP(hipMalloc(&c.shell_order, sizeof(int) * std::max(nb, 1)));
    PTH_HIP(hipMemcpy(c.shell_order, order.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
  }
  std::vector<std::pair<const double *, const double *>> ranges;
  const std::int64_t cells = topo.fine.cells();
  for (size_t f = 0; f < fields.size(); ++f) {
    if (!(fields[f].flags & PTH_FILL_GHOST)) continue;
    const FieldStore &st = md.store[f];
    ranges.emplace_back(st.data, st.data + st.ndata);
    const int nc = fields[f].ncomp;
    for (int v = 0; v < nc; ++v) c.shell.push_back({st.data + v * cells, nc * cells});
  }
  std::vector<Region> sregs, rregs;
  for (size_t q = 0; q < c.pack.host.size(); ++q) {
    const Region &sr = c.pack.host[q], &rr = c.unpack.host[q];
    if (sr.b != rr.b) {
      SetError("internal: send and receive regions do not pair up");
      return PTH_FAIL;
    }
    bool same = false;
    for (const auto &r : ranges)
      same = same || (sr.a >= r.first && sr.a < r.second && rr.a >= r.first && rr.a < r.second);
    if (same) continue;
    sregs.push_back(sr);
    rregs.push_back(rr);
  }
  for (Table *t : {&c.pack_cross, &c.unpack_cross}) {
    t->host = t == &c.pack_cross ? sregs : rregs;
    if (t->Upload() != PTH_COMPLETE) return PTH_FAIL;
  }
  if (c.pack_cross.total != c.unpack_cross.total) {
    SetError("internal: send and receive element counts differ");
    return PTH_FAIL;
  }
  c.cross_built = true;
  *ok = true;
  return PTH_COMPLETE;
}

// Boxes of the physical boundary conditions (reference ApplyBoundaryConditionsOnCoarseOrFine,
// boundary_conditions.cpp:36-58, with GenericBC, boundary_conditions_generic.hpp:174-255):
// per block, FillGhost field and array (data; coarse buffer on multilevel meshes), the brick
// is cut into 3 x 3 x 3 classes (inner ghosts, interior, outer ghosts per dimension); each
// class touching a physical face becomes one box whose cells all map to their source the same
// way. The reference applies the faces one after another over the entire transverse range
// (par_for_bndry over IndexDomain::inner_xD / outer_xD, domain.hpp:184-252); a cell in the
// ghosts of several physical faces therefore ends with the value of the cell reached by
// mapping each of those dimensions, which is what the box's source map gives directly.
pth_status Mesh::BuildBoundaryConditions(MeshData &md, BoundCache &c) {
  c.bc[0].host.clear();
  c.bc[1].host.clear();
  const int nb = static_cast<int>(topo.blocks.size());
  for (int b = 0; b < nb; ++b) {
    const Block &B = topo.blocks[b];
    bool phys[6];
    bool any = false;
    for (int face = 0; face < 6; ++face) {
      const int d = face / 2;
      const std::int64_t last = (topo.nbase[d] << B.at.level) - 1;
      phys[face] = topo.active[d] && topo.bc[face] != PTH_BC_PERIODIC &&
                   (face % 2 == 0 ? B.at.x[d] == 0 : B.at.x[d] == last);
      any = any || phys[face];
    }
    if (!any) continue;
    for (size_t f = 0; f < fields.size(); ++f) {
      if (!(fields[f].flags & PTH_FILL_GHOST)) continue;
      const int nc = fields[f].ncomp;
      for (int coarse = 0; coarse < (topo.multilevel ? 2 : 1); ++coarse) {
        const Extent &e = coarse ? topo.coarse : topo.fine;
        double *base = coarse ? md.store[f].coarse : md.store[f].data;
        if (!base) continue;
        double *slab = base + static_cast<std::int64_t>(b) * nc * e.cells();
        for (int cz = 0; cz < 3; ++cz)
          for (int cy = 0; cy < 3; ++cy)
            for (int cx = 0; cx < 3; ++cx) {
              const int cls[3] = {cx, cy, cz};
              BcBox x;
              std::memset(&x, 0, sizeof(x));
              x.slab = slab;
              x.ncomp = nc;
              x.vec = (fields[f].flags & PTH_VECTOR) ? 1 : 0;
              bool mapped = false, empty = false;
              for (int d = 0; d < 3; ++d) {
                const int lo = e.lo(d), hi = e.hi(d), ent = e.ent(d);
                const int r0 = cls[d] == 0 ? 0 : cls[d] == 1 ? lo : hi + 1;
                const int r1 = cls[d] == 0 ? lo - 1 : cls[d] == 1 ? hi : ent - 1;
                if (r1 < r0) empty = true;
                x.lo[d] = r0;
                x.n[d] = r1 - r0 + 1;
                x.e[d] = ent;
                x.a[d] = 0;
                x.b[d] = 1;
                const int face = cls[d] == 0 ? 2 * d : cls[d] == 2 ? 2 * d + 1 : -1;
                if (face < 0 || !phys[face]) continue;
                mapped = true;
                const int ref = cls[d] == 0 ? lo : hi;
                if (topo.bc[face] == PTH_BC_REFLECT || topo.bc[face] == PTH_BC_FIXED_FACE) {
                  x.a[d] = 2 * ref + (cls[d] == 0 ? -1 : 1);
                  x.b[d] = -1;
                  if (topo.bc[face] == PTH_BC_REFLECT)
                    x.flip |= 1 << d;
                  else
                    x.fixed |= 1 << d;
                } else {
                  x.a[d] = ref;
                  x.b[d] = 0;
                }
              }
              if (empty || !mapped) continue;
              // every written and every read cell lies inside the brick
              for (int d = 0; d < 3; ++d) {
                const int w0 = x.lo[d], w1 = x.lo[d] + x.n[d] - 1;
                const int s0 = x.a[d] + x.b[d] * w0, s1 = x.a[d] + x.b[d] * w1;
                if (w0 < 0 || w1 >= x.e[d] || std::min(s0, s1) < 0 ||
                    std::max(s0, s1) >= x.e[d]) {
                  SetError("internal: a boundary-condition box exceeds its brick");
                  return PTH_FAIL;
                }
              }
              c.bc[coarse].host.push_back(x);
            }
      }
    }
  }
  for (BcTable *t : {&c.bc[0], &c.bc[1]})
    if (t->Upload() != PTH_COMPLETE) return PTH_FAIL;
  return PTH_COMPLETE;
}

void FluxCorrection::Free() {
  if (dev) (void)hipFree(dev);
  if (dev_face) (void)hipFree(dev_face);
  if (store) (void)hipFree(store);
  dev = nullptr;
  dev_face = nullptr;
  store = nullptr;
  built = false;
}

// Geometry of the flux-correction patch that block `fine` sends to its coarser face neighbour
// `coarse` lying across direction d on side sgn (+1 above, -1 below): the fine face index, the
// receiver's face index, and the transverse ranges (the fine block covers the half of the
// receiver's face given by the parity of its own lattice position).
FluxPatch PatchGeometry(const Topology &topo, int fine, int coarse, int d, int sgn) {
  const Extent &fe = topo.fine;
  const Block &B = topo.blocks[fine];
  const int t[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
  FluxPatch p;
  std::memset(&p, 0, sizeof(p));
  p.fine = fine;
  p.coarse = coarse;
  p.d = d;
  p.fface = sgn > 0 ? fe.lo(d) + fe.n[d] : fe.lo(d);
  p.cface = sgn > 0 ? fe.lo(d) : fe.lo(d) + fe.n[d];
  for (int q = 0; q < 2; ++q) {
    const int dim = t[q];
    const bool act = topo.active[dim];
    p.step[q] = act ? 2 : 1;
    p.n[q] = act ? fe.n[dim] / 2 : 1;
    p.ft[q] = fe.lo(dim);
    p.ct[q] = fe.lo(dim) + (act ? static_cast<int>(B.at.x[dim] & 1) * (fe.n[dim] / 2) : 0);
  }
  return p;
}

// Patches of the flux-correction exchange of field f. A block sends to a face neighbour one
// level coarser (loop_utils.hpp:147-151 with flxcor_send); the receiver's face is then covered
// by 2^(ndim-1) such patches, one per child of its same-level neighbour, which the 2:1 level
// balance of the mesh guarantees and the builder checks.
FluxCorrection *Mesh::FluxPatches(int f) {
  if (f < 0 || f >= static_cast<int>(fields.size()) || !(fields[f].flags & PTH_WITH_FLUXES)) {
    SetError("flux correction needs a field with fluxes");
    return nullptr;
  }
  FluxCorrection &c = flxcor[f];
  if (c.built) return &c;
  const Extent &fe = topo.fine;
  const int nb = static_cast<int>(topo.blocks.size());
  const int V = fields[f].ncomp;
  c.ncomp = V;
  c.host.clear();
  c.face.assign(6 * static_cast<size_t>(nb), -1);
  c.max_cells = 0;
  std::vector<int> hits(6 * static_cast<size_t>(nb), 0);
  std::int64_t total = 0;
  for (int b = 0; b < nb; ++b) {
    const Block &B = topo.blocks[b];
    for (const Link &l : B.links) {
      int nz = 0, d = -1;
      for (int q = 0; q < 3; ++q)
        if (l.off[q] != 0) {
          ++nz;
          d = q;
        }
      if (nz != 1 || l.level != B.at.level - 1) continue;
      const int t[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
      FluxPatch p = PatchGeometry(topo, b, l.gid, d, l.off[d]);
      const int side = l.off[d] > 0 ? 0 : 1;  // receiver's low face when it lies above
      const size_t key = 6 * static_cast<size_t>(l.gid) + 2 * d + side;
      if (c.face[key] < 0) {
        c.face[key] = static_cast<int>(total);
        total += static_cast<std::int64_t>(fe.n[t[0]]) * fe.n[t[1]] * V;
        if (total >= (std::int64_t(1) << 31)) {
          SetError("flux correction store exceeds 2^31 values");
          return nullptr;
        }
      }
      ++hits[key];
      p.rows = fe.n[t[0]];
      p.comp = fe.n[t[0]] * fe.n[t[1]];
      p.dst = c.face[key] + (p.ct[0] - fe.lo(t[0])) + (p.ct[1] - fe.lo(t[1])) * p.rows;
      p.area = B.area[d];
      c.max_cells = std::max(c.max_cells, p.n[0] * p.n[1]);
      c.host.push_back(p);
    }
  }
  for (int b = 0; b < nb; ++b)
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < 2; ++s) {
        const size_t key = 6 * static_cast<size_t>(b) + 2 * d + s;
        if (c.face[key] < 0) continue;
        int want = 1;
        for (int q = 0; q < 3; ++q)
          if (q != d && topo.active[q]) want *= 2;
        if (hits[key] != want) {
          SetError("flux correction: a coarse face is not covered by its finer neighbours");
          return nullptr;
        }
      }
  c.store_n = total;
  if (!c.host.empty()) {
    if (hipMalloc(&c.dev, sizeof(FluxPatch) * c.host.size()) != hipSuccess ||
        hipMalloc(&c.dev_face, sizeof(int) * c.face.size()) != hipSuccess ||
        hipMalloc(&c.store, sizeof(double) * total) != hipSuccess ||
        hipMemset(c.store, 0, sizeof(double) * total) != hipSuccess ||
        hipMemcpy(c.dev, c.host.data(), sizeof(FluxPatch) * c.host.size(),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.dev_face, c.face.data(), sizeof(int) * c.face.size(),
                  hipMemcpyHostToDevice) != hipSuccess) {
      SetError("flux correction: device allocation failed");
      c.Free();
      return nullptr;
    }
  }
  c.built = true;
  return &c;
}

}  // namespace pth

namespace {

int Bt(int bt) { return bt == PTH_BT_LOCAL ? PTH_BT_ANY : bt; }  // same channel group

pth_status NeedFinal(const Mesh *m) {
  if (!m || !m->finalized) {
    SetError("mesh is not finalized");
    return PTH_FAIL;
  }
  return PTH_COMPLETE;
}

// SplitMix64 (Steele, Lea, Flood 2014): the seeded generator of the synthetic inputs
struct SplitMix64 {
  std::uint64_t s;
  std::uint64_t next() {
    std::uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

}  // namespace


extern "C" {

const char *pth_last_error(void) { return ErrorText(); }
const char *pth_version(void) { return "parthenon_amd 0.2 (gfx950)"; }

pth_status pth_device_memory(int64_t *free_bytes, int64_t *total_bytes) {
  size_t f = 0, t = 0;
  PTH_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = static_cast<int64_t>(f);
  if (total_bytes) *total_bytes = static_cast<int64_t>(t);
  return PTH_COMPLETE;
}

pth_mesh *pth_mesh_create(const int nx[3], const int bnx[3], int nghost, const double xmin[3],
                          const double xmax[3], const int periodic[3]) {
  Mesh *m = new Mesh;
  const std::string err = m->topo.Init(nx, bnx, nghost, xmin, xmax, periodic);
  if (!err.empty()) {
    SetError(err);
    delete m;
    return nullptr;
  }
  return reinterpret_cast<pth_mesh *>(m);
}

void pth_mesh_destroy(pth_mesh *m) { delete Impl(m); }

pth_status pth_mesh_set_boundary(pth_mesh *mh, int face, int type) {
  Mesh *m = Impl(mh);
  if (m->planned) {
    SetError("pth_mesh_set_boundary after pth_mesh_plan / pth_mesh_finalize");
    return PTH_FAIL;
  }
  if (face < 0 || face >= 6) {
    SetError("boundary face must be 0..5 (inner x1, outer x1, ..., outer x3)");
    return PTH_FAIL;
  }
  if (type != PTH_BC_OUTFLOW && type != PTH_BC_REFLECT && type != PTH_BC_FIXED_FACE) {
    SetError("boundary condition must be PTH_BC_OUTFLOW, PTH_BC_REFLECT or PTH_BC_FIXED_FACE");
    return PTH_FAIL;
  }
  if (m->topo.periodic[face / 2]) {
    SetError("the faces of a periodic dimension stay periodic");
    return PTH_FAIL;
  }
  m->topo.bc[face] = type;
  return PTH_COMPLETE;
}

pth_status pth_mesh_refine_box(pth_mesh *mh, const double box[6], int level) {
  Mesh *m = Impl(mh);
  if (m->planned || level < 1) {
    SetError("refinement must precede finalize and have level >= 1");
    return PTH_FAIL;
  }
  m->topo.RefineBox(box, level);
  return PTH_COMPLETE;
}

pth_status pth_mesh_refine_sphere(pth_mesh *mh, const double center[3], double radius,
                                  int level) {
  Mesh *m = Impl(mh);
  if (m->planned || level < 1) {
    SetError("refinement must precede finalize and have level >= 1");
    return PTH_FAIL;
  }
  m->topo.RefineSphere(center, radius, level);
  return PTH_COMPLETE;
}

pth_status pth_mesh_add_field(pth_mesh *mh, const char *name, int ncomp, unsigned flags) {
  Mesh *m = Impl(mh);
  if (m->finalized) {
    SetError("fields must be added before finalize");
    return PTH_FAIL;
  }
  if (ncomp < 1 || m->FieldIndex(name) >= 0) {
    SetError(std::string("bad or duplicate field ") + name);
    return PTH_FAIL;
  }
  m->fields.push_back(FieldSpec{name, ncomp, flags});
  return PTH_COMPLETE;
}

pth_status pth_mesh_register_refinement_ops(pth_mesh *mh, const char *field, int prolong,
                                            int restrict_op) {
  Mesh *m = Impl(mh);
  if (m->finalized) {
    SetError("refinement operators must be registered before finalize");
    return PTH_FAIL;
  }
  const int f = field ? m->FieldIndex(field) : -1;
  if (f < 0) {
    SetError(std::string("no field ") + (field ? field : "(null)"));
    return PTH_FAIL;
  }
  if ((prolong != PTH_PROLONG_MINMOD && prolong != PTH_PROLONG_LINEAR) ||
      restrict_op != PTH_RESTRICT_AVERAGE) {
    SetError("unknown refinement operator");
    return PTH_FAIL;
  }
  m->fields[f].prolong = prolong;
  return PTH_COMPLETE;
}

pth_status pth_mesh_plan(pth_mesh *mh) { return Impl(mh)->Plan(); }

pth_status pth_mesh_finalize(pth_mesh *mh) { return Impl(mh)->Finalize(); }

int pth_mesh_neighbors(const pth_mesh *mh, int gid, int64_t *out, int max) {
  const Mesh *m = Impl(mh);
  if (!m->planned || gid < 0 || gid >= static_cast<int>(m->topo.blocks.size())) return -1;
  int n = 0;
  for (const Link &l : m->topo.blocks[gid].links) {
    if (n >= max) break;
    int64_t *o = out + 8 * n;
    o[0] = l.gid;
    o[1] = l.level;
    for (int d = 0; d < 3; ++d) {
      o[2 + d] = l.seen[d];
      o[5 + d] = l.off[d];
    }
    ++n;
  }
  return n;
}

int pth_mesh_region_boxes(const pth_mesh *mh, int gid, int *out, int max) {
  const Mesh *m = Impl(mh);
  if (!m->planned || gid < 0 || gid >= static_cast<int>(m->topo.blocks.size())) return -1;
  const Block &B = m->topo.blocks[gid];
  int n = 0;
  auto put = [&](int kind, const Box &b) {
    if (n >= max) return;
    int *o = out + 7 * n;
    o[0] = kind;
    for (int d = 0; d < 3; ++d) {
      o[1 + 2 * d] = b.lo[d];
      o[2 + 2 * d] = b.hi[d];
    }
    ++n;
  };
  for (const Link &l : B.links) {
    const bool coarser = l.level < B.at.level;
    put(0, m->topo.Indices(B, l, l.is_cell() ? Side::kSendInterior : Side::kSendBoundary, false));
    put(1, m->topo.Indices(B, l, l.is_cell() ? Side::kRecvInterior : Side::kRecvBoundary, false));
    if (!m->topo.multilevel) continue;
    if (coarser) {
      put(2, m->topo.Indices(B, l, Side::kSendBoundary, true));
      put(3, m->topo.Indices(B, l, Side::kRecvBoundary, true));
    } else if (B.coarser_neighbor && B.at.level > 0) {
      put(4, m->topo.Indices(B, l, Side::kRecvBoundary, true));
    }
    // flux correction: 5 restrict+send box (coarse index space), 6 receive box
    int nz = 0, d = -1;
    for (int q = 0; q < 3; ++q)
      if (l.off[q] != 0) {
        ++nz;
        d = q;
      }
    if (nz != 1 || (l.level != B.at.level - 1 && l.level != B.at.level + 1)) continue;
    const int t[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
    Box box;
    if (l.level == B.at.level - 1) {
      const FluxPatch p = PatchGeometry(m->topo, gid, l.gid, d, l.off[d]);
      const Extent &ce = m->topo.coarse;
      box.lo[d] = box.hi[d] = ce.lo(d) + (p.fface == m->topo.fine.lo(d) ? 0 : ce.n[d]);
      for (int q = 0; q < 2; ++q) {
        box.lo[t[q]] = ce.lo(t[q]);
        box.hi[t[q]] = ce.lo(t[q]) + p.n[q] - 1;
      }
      put(5, box);
    } else {
      const FluxPatch p = PatchGeometry(m->topo, l.gid, gid, d, -l.off[d]);
      box.lo[d] = box.hi[d] = p.cface;
      for (int q = 0; q < 2; ++q) {
        box.lo[t[q]] = p.ct[q];
        box.hi[t[q]] = p.ct[q] + p.n[q] - 1;
      }
      put(6, box);
    }
  }
  return n;
}

pth_status pth_mesh_set_stream(pth_mesh *mh, void *s) {
  Impl(mh)->stream = static_cast<hipStream_t>(s);
  return PTH_COMPLETE;
}

pth_status pth_mesh_synchronize(pth_mesh *mh) {
  PTH_HIP(hipStreamSynchronize(Impl(mh)->stream));
  return PTH_COMPLETE;
}

pth_status pth_mesh_set_pack_mode(pth_mesh *mh, int mode) {
  if (mode != PTH_PACK_PERSISTENT && mode != PTH_PACK_PER_BUFFER) {
    SetError("unknown pack mode");
    return PTH_FAIL;
  }
  Impl(mh)->pack_mode = mode;
  return PTH_COMPLETE;
}

int pth_mesh_nblocks(const pth_mesh *mh) {
  if (!Impl(mh)->planned) return -1;
  return static_cast<int>(Impl(mh)->topo.blocks.size());
}

int pth_mesh_multilevel(const pth_mesh *mh) { return Impl(mh)->topo.multilevel ? 1 : 0; }

pth_status pth_mesh_block_info(const pth_mesh *mh, int gid, int64_t out[4], double x[6]) {
  const Mesh *m = Impl(mh);
  if (!m->planned) {
    SetError("mesh is not planned");
    return PTH_FAIL;
  }
  if (gid < 0 || gid >= static_cast<int>(m->topo.blocks.size())) {
    SetError("block id out of range");
    return PTH_FAIL;
  }
  const Block &b = m->topo.blocks[gid];
  out[0] = b.at.level;
  for (int d = 0; d < 3; ++d) {
    out[1 + d] = b.at.x[d];
    x[d] = b.xmin[d];
    x[3 + d] = b.xmax[d];
  }
  return PTH_COMPLETE;
}

pth_status pth_mesh_index_shapes(const pth_mesh *mh, int out[12]) {
  const Mesh *m = Impl(mh);
  if (!m->planned) {
    SetError("mesh is not planned");
    return PTH_FAIL;
  }
  for (int d = 0; d < 3; ++d) {
    out[d] = m->topo.fine.n[d];
    out[3 + d] = m->topo.fine.g[d];
    out[6 + d] = m->topo.coarse.n[d];
    out[9 + d] = m->topo.coarse.g[d];
  }
  return PTH_COMPLETE;
}

pth_status pth_mesh_exchange_counts(const pth_mesh *mh, int bt, int64_t *nbuffers,
                                    int64_t *nelements) {
  const Mesh *m = Impl(mh);
  if (!m->planned) {
    SetError("mesh is not planned");
    return PTH_FAIL;
  }
  std::int64_t nbuf = 0, nel = 0;
  if (Bt(bt) == PTH_BT_ANY) {
    for (const Block &B : m->topo.blocks)
      for (const FieldSpec &f : m->fields) {
        if (!(f.flags & PTH_FILL_GHOST)) continue;
        for (const Link &l : B.links) {
          const Side s = l.is_cell() ? Side::kSendInterior : Side::kSendBoundary;
          nel += m->topo.Indices(B, l, s, false).count() * f.ncomp;
          ++nbuf;
        }
      }
  }
  *nbuffers = nbuf;
  *nelements = nel;
  return PTH_COMPLETE;
}

pth_meshdata *pth_meshdata_get(pth_mesh *mh, const char *name) {
  Mesh *m = Impl(mh);
  if (NeedFinal(m) != PTH_COMPLETE) return nullptr;
  return reinterpret_cast<pth_meshdata *>(m->Get(name));
}

pth_status pth_field_pointers(pth_meshdata *mdh, const char *field, double **data,
                              double **coarse, double **flux) {
  MeshData *md = Impl(mdh);
  const int f = md->mesh->FieldIndex(field);
  if (f < 0) {
    SetError(std::string("no field ") + field);
    return PTH_FAIL;
  }
  if (data) *data = md->store[f].data;
  if (coarse) *coarse = md->store[f].coarse;
  if (flux) *flux = md->store[f].flux;
  return PTH_COMPLETE;
}

// ---- packs ------------------------------------------------------------------------------
pth_pack *pth_pack_get(pth_meshdata *mdh, const char *const *fields, int nfields,
                       unsigned options) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (!fields || nfields < 1) {
    SetError("pth_pack_get: no fields");
    return nullptr;
  }
  if (options & ~static_cast<unsigned>(PTH_PACK_WITH_FLUXES | PTH_PACK_COARSE |
                                       PTH_PACK_FLATTEN)) {
    SetError("pth_pack_get: unknown option");
    return nullptr;
  }
  const bool with_flux = options & PTH_PACK_WITH_FLUXES, coarse = options & PTH_PACK_COARSE;
  std::string key = std::to_string(options);
  std::vector<int> fid;
  for (int q = 0; q < nfields; ++q) {
    const int f = fields[q] ? m->FieldIndex(fields[q]) : -1;
    if (f < 0) {
      SetError(std::string("pth_pack_get: no field ") + (fields[q] ? fields[q] : "(null)"));
      return nullptr;
    }
    if (with_flux && !md->store[f].flux) {
      SetError(std::string("pth_pack_get: field ") + fields[q] + " has no fluxes");
      return nullptr;
    }
    if (coarse && !md->store[f].coarse) {
      SetError(std::string("pth_pack_get: field ") + fields[q] +
               " has no coarse buffer (single-level mesh)");
      return nullptr;
    }
    fid.push_back(f);
    key += "|";
    key += fields[q];
  }
  PackEntry &e = md->packs[key];
  if (e.dev && e.generation == m->generation)
    return reinterpret_cast<pth_pack *>(&e);
  e.Free();
  const int nb = static_cast<int>(m->topo.blocks.size());
  const Extent &x = coarse ? m->topo.coarse : m->topo.fine;
  const std::int64_t cells = x.cells();
  int nvar = 0;
  e.bounds.assign(2 * nfields, 0);
  for (int q = 0; q < nfields; ++q) {
    e.bounds[q] = nvar;
    nvar += m->fields[fid[q]].ncomp;
    e.bounds[nfields + q] = nvar - 1;
  }
  std::vector<double *> ptr(static_cast<size_t>(nb) * nvar * (with_flux ? 4 : 1));
  for (int b = 0; b < nb; ++b)
    for (int q = 0; q < nfields; ++q) {
      const FieldStore &fs = md->store[fid[q]];
      const int nc = m->fields[fid[q]].ncomp;
      for (int c = 0; c < nc; ++c) {
        const std::int64_t off = (static_cast<std::int64_t>(b) * nc + c) * cells;
        const int v = e.bounds[q] + c;
        ptr[static_cast<size_t>(b) * nvar + v] = (coarse ? fs.coarse : fs.data) + off;
        if (with_flux)
          for (int d = 0; d < 3; ++d)
            ptr[static_cast<size_t>(nb) * nvar * (1 + d) + static_cast<size_t>(b) * nvar + v] =
                fs.flux + d * fs.ndata + off;
      }
    }
  const size_t pbytes = ptr.size() * sizeof(double *);
  const size_t bbytes = e.bounds.size() * sizeof(int);
  if (hipMalloc(&e.dev, pbytes + bbytes) != hipSuccess ||
      hipMemcpy(e.dev, ptr.data(), pbytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(static_cast<char *>(e.dev) + pbytes, e.bounds.data(), bbytes,
                hipMemcpyHostToDevice) != hipSuccess) {
    SetError("pth_pack_get: device table upload failed");
    e.Free();
    return nullptr;
  }
  pth_pack_view &v = e.view;
  v.data = static_cast<double *const *>(e.dev);
  v.flux = with_flux ? static_cast<double *const *>(e.dev) + static_cast<size_t>(nb) * nvar
                     : nullptr;
  v.bounds = reinterpret_cast<const int *>(static_cast<char *>(e.dev) + pbytes);
  v.coords = reinterpret_cast<const double *>(m->coef);
  v.geom = reinterpret_cast<const double *>(m->geom);
  v.nblocks = nb;
  v.nvar = nvar;
  v.nfields = nfields;
  v.options = static_cast<int>(options);
  for (int d = 0; d < 3; ++d) v.ent[d] = x.ent(d);
  v.is = x.lo(0);
  v.ie = x.hi(0);
  v.js = x.lo(1);
  v.je = x.hi(1);
  v.ks = x.lo(2);
  v.ke = x.hi(2);
  e.generation = m->generation;
  return reinterpret_cast<pth_pack *>(&e);
}

pth_status pth_pack_view_get(const pth_pack *p, pth_pack_view *view) {
  if (!p || !view) {
    SetError("pth_pack_view_get: null argument");
    return PTH_FAIL;
  }
  *view = reinterpret_cast<const PackEntry *>(p)->view;
  return PTH_COMPLETE;
}

pth_status pth_pack_bounds(const pth_pack *p, int field, int *lo, int *hi) {
  const PackEntry *e = reinterpret_cast<const PackEntry *>(p);
  if (!e || field < 0 || field >= e->view.nfields) {
    SetError("pth_pack_bounds: no such field in the pack");
    return PTH_FAIL;
  }
  if (lo) *lo = e->bounds[field];
  if (hi) *hi = e->bounds[e->view.nfields + field];
  return PTH_COMPLETE;
}

static pth_status FieldBuffer(MeshData *md, const char *field, int which, double **p,
                              std::int64_t *n) {
  const int f = md->mesh->FieldIndex(field);
  if (f < 0) {
    SetError(std::string("no field ") + field);
    return PTH_FAIL;
  }
  FieldStore &s = md->store[f];
  if (which == 0) {
    *p = s.data;
    *n = s.ndata;
  } else if (which == 1) {
    *p = s.coarse;
    *n = s.ncoarse;
  } else if (which == 2) {
    *p = s.flux;
    *n = s.nflux;
  } else {
    SetError("which must be 0 (data), 1 (coarse) or 2 (flux)");
    return PTH_FAIL;
  }
  return PTH_COMPLETE;
}

pth_status pth_field_upload(pth_meshdata *mdh, const char *field, int which, const double *host,
                            size_t n) {
  MeshData *md = Impl(mdh);
  double *p;
  std::int64_t cap;
  if (FieldBuffer(md, field, which, &p, &cap) != PTH_COMPLETE) return PTH_FAIL;
  if (static_cast<std::int64_t>(n) != cap) {
    SetError("size mismatch in pth_field_upload");
    return PTH_FAIL;
  }
  PTH_HIP(hipMemcpyAsync(p, host, sizeof(double) * n, hipMemcpyHostToDevice, md->mesh->stream));
  PTH_HIP(hipStreamSynchronize(md->mesh->stream));
  return PTH_COMPLETE;
}

pth_status pth_field_download(pth_meshdata *mdh, const char *field, int which, double *host,
                              size_t n) {
  MeshData *md = Impl(mdh);
  double *p;
  std::int64_t cap;
  if (FieldBuffer(md, field, which, &p, &cap) != PTH_COMPLETE) return PTH_FAIL;
  if (static_cast<std::int64_t>(n) != cap) {
    SetError("size mismatch in pth_field_download");
    return PTH_FAIL;
  }
  PTH_HIP(hipMemcpyAsync(host, p, sizeof(double) * n, hipMemcpyDeviceToHost, md->mesh->stream));
  PTH_HIP(hipStreamSynchronize(md->mesh->stream));
  return PTH_COMPLETE;
}

// ---- boundary communication ----

static bool IsFlxcor(int bt) { return bt == PTH_BT_FLXCOR_SEND || bt == PTH_BT_FLXCOR_RECV; }

// Flux correction over every field with fluxes (SendBoundBufs<flxcor_send>, ReceiveBoundBufs /
// SetBounds<flxcor_recv>): the restricted fine fluxes go to the correction store, then into
// the coarse blocks' stored fluxes. Nothing to do on a single-level mesh (no channels).
static pth_status FlxcorStep(MeshData *md, int step) {
  Mesh *m = md->mesh;
  if (NeedFinal(m) != PTH_COMPLETE) return PTH_FAIL;
  if (!m->topo.multilevel) return PTH_COMPLETE;
  if (step == 1) {  // receive: one GPU, the store is ordered on the same stream
    if (m->flxcor_state == BufState::kSending) m->flxcor_state = BufState::kReceived;
    return m->flxcor_state == BufState::kReceived ? PTH_COMPLETE : PTH_INCOMPLETE;
  }
  if (step == 0 && m->flxcor_state != BufState::kStale) return PTH_INCOMPLETE;
  if (step == 2 && m->flxcor_state != BufState::kReceived) {
    SetError("SetBounds: flux corrections are not in a received state");
    return PTH_FAIL;
  }
  for (size_t f = 0; f < m->fields.size(); ++f) {
    if (!(m->fields[f].flags & PTH_WITH_FLUXES)) continue;
    FluxCorrection *fc = m->FluxPatches(static_cast<int>(f));
    if (!fc) return PTH_FAIL;
    const pth_status st = step == 0 ? FlxcorSend(m->topo, *fc, md->store[f].flux, m->stream)
                                    : FlxcorSet(m->topo, *fc, md->store[f].flux, m->stream);
    if (st != PTH_COMPLETE) return PTH_FAIL;
  }
  m->flxcor_state = step == 0 ? BufState::kSending : BufState::kStale;
  return PTH_COMPLETE;
}

pth_status pth_send_bound_bufs(pth_meshdata *mdh, int bt) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (IsFlxcor(bt)) return FlxcorStep(md, 0);
  bt = Bt(bt);
  if (m->BuildCache(*md, bt) != PTH_COMPLETE) return PTH_FAIL;
  BoundCache &c = md->cache[bt];
  if (c.nbuffers == 0) return PTH_COMPLETE;
  if (m->state[bt] != BufState::kStale) return PTH_INCOMPLETE;  // previous exchange unfinished
  if (m->topo.multilevel &&
      LaunchRestrictT(c.restrict_send, m->topo, m->geom, m->stream) != PTH_COMPLETE)
    return PTH_FAIL;
  if (LaunchPack(c.pack, nullptr, m->stream, true, m->pack_mode) != PTH_COMPLETE)
    return PTH_FAIL;
  m->state[bt] = BufState::kSending;
  return PTH_COMPLETE;
}

pth_status pth_start_receive_bound_bufs(pth_meshdata *mdh, int bt) {
  MeshData *md = Impl(mdh);
  if (IsFlxcor(bt)) return NeedFinal(md->mesh);
  return md->mesh->BuildCache(*md, Bt(bt));
}

pth_status pth_receive_bound_bufs(pth_meshdata *mdh, int bt) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (IsFlxcor(bt)) return FlxcorStep(md, 1);
  bt = Bt(bt);
  if (m->BuildCache(*md, bt) != PTH_COMPLETE) return PTH_FAIL;
  if (md->cache[bt].nbuffers == 0) return PTH_COMPLETE;
  // one GPU: a sent buffer is received as soon as its pack is ordered before the unpack on
  // the same stream
  if (m->state[bt] == BufState::kSending) m->state[bt] = BufState::kReceived;
  return m->state[bt] == BufState::kReceived ? PTH_COMPLETE : PTH_INCOMPLETE;
}

pth_status pth_set_bounds(pth_meshdata *mdh, int bt) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (IsFlxcor(bt)) return FlxcorStep(md, 2);
  bt = Bt(bt);
  if (m->BuildCache(*md, bt) != PTH_COMPLETE) return PTH_FAIL;
  BoundCache &c = md->cache[bt];
  if (c.nbuffers == 0) return PTH_COMPLETE;
  if (m->state[bt] != BufState::kReceived) {
    SetError("SetBounds: buffers are not in a received state");
    return PTH_FAIL;
  }
  if (LaunchPack(c.unpack, nullptr, m->stream, false, m->pack_mode) != PTH_COMPLETE)
    return PTH_FAIL;
  m->state[bt] = BufState::kStale;
  if (m->topo.multilevel &&
      LaunchRestrictT(c.restrict_set, m->topo, m->geom, m->stream) != PTH_COMPLETE)
    return PTH_FAIL;
  return PTH_COMPLETE;
}

pth_status pth_apply_boundary_conditions(pth_meshdata *mdh, int coarse) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (NeedFinal(m) != PTH_COMPLETE) return PTH_FAIL;
  if (m->BuildCache(*md, PTH_BT_ANY) != PTH_COMPLETE) return PTH_FAIL;
  return LaunchBoundaryConditions(md->cache[PTH_BT_ANY].bc[coarse ? 1 : 0], m->stream);
}

pth_status pth_prolongate_bounds(pth_meshdata *mdh, int bt) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  bt = Bt(bt);
  if (m->BuildCache(*md, bt) != PTH_COMPLETE) return PTH_FAIL;
  BoundCache &c = md->cache[bt];
  if (c.nbuffers == 0 || !m->topo.multilevel) return PTH_COMPLETE;
  if (LaunchProlongT(c.prolong, m->topo, m->geom, m->stream, true) != PTH_COMPLETE)
    return PTH_FAIL;
  return LaunchProlongT(c.prolong_linear, m->topo, m->geom, m->stream, false);
}

// ---- Burgers package and Update:: tasks ----

static int NeedField(MeshData *md, const char *name, int min_comp) {
  const int f = md->mesh->FieldIndex(name);
  if (f < 0 || md->mesh->fields[f].ncomp < min_comp) {
    SetError(std::string("Burgers tasks need field ") + name);
    return -1;
  }
  return f;
}

pth_status pth_burgers_calculate_fluxes(pth_meshdata *mdh, int recon) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int f = NeedField(md, "U", 4);
  if (f < 0) return PTH_FAIL;
  if (!(m->fields[f].flags & PTH_WITH_FLUXES)) {
    SetError("field U needs WithFluxes");
    return PTH_FAIL;
  }
  // burgers_package::Initialize (burgers_package.cpp:51-61): WENO5 needs nghost >= 4; the
  // linear reconstruction reads one cell beyond the extended box and needs nghost >= 2
  const int need = recon == 0 ? 4 : 2;
  if (recon != 0 && recon != 1) {
    SetError("recon must be 0 (weno5) or 1 (linear)");
    return PTH_FAIL;
  }
  for (int d = 0; d < m->topo.ndim; ++d)
    if (m->topo.fine.g[d] < need) {
      SetError(recon == 0 ? "weno5 reconstruction requires 4 or more ghost cells"
                          : "linear reconstruction requires 2 or more ghost cells");
      return PTH_FAIL;
    }
  const int V = m->fields[f].ncomp;
  if (!m->face || m->face_V != V) {
    if (m->face) (void)hipFree(m->face);
    m->face = nullptr;
    if (Alloc(&m->face, 6 * md->store[f].ndata) != PTH_COMPLETE) return PTH_FAIL;
    m->face_V = V;
  }
  if (BurgersRecon(m->topo, V, md->store[f].data, m->face, recon, m->stream) != PTH_COMPLETE)
    return PTH_FAIL;
  return BurgersFlux(m->topo, V, m->face, md->store[f].flux, m->stream);
}

pth_status pth_flux_divergence(pth_meshdata *inh, pth_meshdata *dudth) {
  MeshData *in = Impl(inh), *out = Impl(dudth);
  Mesh *m = in->mesh;
  for (size_t f = 0; f < m->fields.size(); ++f) {
    if (!(m->fields[f].flags & PTH_WITH_FLUXES)) continue;
    if (FluxDivergence(m->topo, m->fields[f].ncomp, in->store[f].flux, out->store[f].data,
                       m->coef, m->stream) != PTH_COMPLETE)
      return PTH_FAIL;
  }
  return PTH_COMPLETE;
}

pth_status pth_weighted_sum_independent(pth_meshdata *in1h, pth_meshdata *in2h, double w1,
                                        double w2, pth_meshdata *outh) {
  MeshData *x = Impl(in1h), *y = Impl(in2h), *z = Impl(outh);
  Mesh *m = x->mesh;
  for (size_t f = 0; f < m->fields.size(); ++f) {
    if (!(m->fields[f].flags & PTH_INDEPENDENT)) continue;
    if (WeightedSum(x->store[f].data, y->store[f].data, w1, w2, z->store[f].data,
                    z->store[f].ndata, m->stream) != PTH_COMPLETE)
      return PTH_FAIL;
  }
  return PTH_COMPLETE;
}

pth_status pth_burgers_calculate_derived(pth_meshdata *mdh) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int fu = NeedField(md, "U", 4), fd = NeedField(md, "derived", 1);
  if (fu < 0 || fd < 0) return PTH_FAIL;
  return BurgersDerived(m->topo, m->fields[fu].ncomp, md->store[fu].data, md->store[fd].data,
                        m->stream);
}

pth_status pth_burgers_estimate_timestep(pth_meshdata *mdh, double cfl, double *dt) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int fu = NeedField(md, "U", 3);
  if (fu < 0) return PTH_FAIL;
  if (BurgersDtMin(m->topo, m->fields[fu].ncomp, md->store[fu].data, m->coef, m->scratch,
                   static_cast<int>(m->scratch_n), m->scratch + m->scratch_n,
                   m->stream) != PTH_COMPLETE)
    return PTH_FAIL;
  PTH_HIP(hipMemcpyAsync(m->pinned, m->scratch + m->scratch_n, sizeof(double),
                         hipMemcpyDeviceToHost, m->stream));
  PTH_HIP(hipStreamSynchronize(m->stream));
  *dt = cfl * m->pinned[0];
  return PTH_COMPLETE;
}

// History: the 8 octant masses of burgers_package (burgers_package.cpp:96-135, 406-439)
pth_status pth_burgers_mass_history(pth_meshdata *mdh, double out[8]) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int fu = NeedField(md, "U", 3);
  if (fu < 0) return PTH_FAIL;
  // partials in the first scratch slots, the 8 results after them
  const int nscr = static_cast<int>(std::min<std::int64_t>(m->scratch_n - 8, 8 * 2048));
  if (BurgersMassHistory(m->topo, m->fields[fu].ncomp, md->store[fu].data, m->geom, m->coef,
                         m->topo.xmin, m->topo.xmax, m->scratch, nscr, m->scratch + nscr,
                         m->stream) != PTH_COMPLETE)
    return PTH_FAIL;
  PTH_HIP(hipMemcpyAsync(m->pinned, m->scratch + nscr, sizeof(double) * 8,
                         hipMemcpyDeviceToHost, m->stream));
  PTH_HIP(hipStreamSynchronize(m->stream));
  for (int q = 0; q < 8; ++q) out[q] = m->pinned[q];
  return PTH_COMPLETE;
}

static pth_status PgenUpload(MeshData *md, int f, const std::vector<double> &h) {
  Mesh *m = md->mesh;
  PTH_HIP(hipMemcpy(md->store[f].data, h.data(), sizeof(double) * h.size(),
                    hipMemcpyHostToDevice));
  (void)m;
  return PTH_COMPLETE;
}

// Seeded sum-of-sinusoids initial condition for U over block interiors: for component c and
// mode m = 0..3 draw A ~ U(0.1, 1) then phi ~ U(0, 2 pi) from SplitMix64(seed) in (c, m)
// order; mode m varies along direction (c + m) % 3 with wavenumber m + 1.
pth_status pth_burgers_pgen_sinusoid(pth_meshdata *mdh, uint64_t seed) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int f = NeedField(md, "U", 3);
  if (f < 0) return PTH_FAIL;
  const int V = m->fields[f].ncomp;
  SplitMix64 rng{seed};
  std::vector<double> A(4 * V), P(4 * V);
  for (int c = 0; c < V; ++c)
    for (int q = 0; q < 4; ++q) {
      A[4 * c + q] = 0.1 + 0.9 * rng.uniform();
      P[4 * c + q] = 2.0 * M_PI * rng.uniform();
    }
  const Extent &e = m->topo.fine;
  std::vector<double> h(md->store[f].ndata, 0.0);
  for (size_t b = 0; b < m->topo.blocks.size(); ++b) {
    const Block &B = m->topo.blocks[b];
    for (int c = 0; c < V; ++c)
      for (int k = e.lo(2); k <= e.hi(2); ++k)
        for (int j = e.lo(1); j <= e.hi(1); ++j)
          for (int i = e.lo(0); i <= e.hi(0); ++i) {
            const double x[3] = {B.x0[0] + (i + 0.5) * B.dx[0], B.x0[1] + (j + 0.5) * B.dx[1],
                                 B.x0[2] + (k + 0.5) * B.dx[2]};
            double v = 0.0;
            for (int q = 0; q < 4; ++q)
              v += A[4 * c + q] * std::sin(2.0 * M_PI * (q + 1) * x[(c + q) % 3] + P[4 * c + q]);
            h[((static_cast<std::int64_t>(b) * V + c) * e.ent(2) + k) * e.ent(1) * e.ent(0) +
              static_cast<std::int64_t>(j) * e.ent(0) + i] = v;
          }
  }
  return PgenUpload(md, f, h);
}

// The reference problem generator (benchmarks/burgers/parthenon_app_inputs.cpp:34-79).
pth_status pth_burgers_pgen_vibe(pth_meshdata *mdh) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int f = NeedField(md, "U", 3);
  if (f < 0) return PTH_FAIL;
  const int V = m->fields[f].ncomp;
  const Extent &e = m->topo.fine;
  std::vector<double> h(md->store[f].ndata, 0.0);
  for (size_t b = 0; b < m->topo.blocks.size(); ++b) {
    const Block &B = m->topo.blocks[b];
    for (int k = e.lo(2); k <= e.hi(2); ++k)
      for (int j = e.lo(1); j <= e.hi(1); ++j)
        for (int i = e.lo(0); i <= e.hi(0); ++i) {
          const double x = B.x0[0] + (i + 0.5) * B.dx[0], y = B.x0[1] + (j + 0.5) * B.dx[1],
                       z = B.x0[2] + (k + 0.5) * B.dx[2];
          const double qv[3] = {(std::tanh(-20.0 * x) * std::cos(M_PI * x) + 1.0) *
                                    std::exp(-30.0 * y * y) * std::exp(-30.0 * z * z),
                                (std::sin(M_PI * y) + 0.2) * std::exp(-30.0 * x * x) *
                                    std::exp(-30.0 * z * z),
                                (std::tanh(-20. * z) * std::cos(M_PI * z) + 0.5) *
                                    std::exp(-30.0 * x * x) * std::exp(-30.0 * y * y)};
          for (int c = 0; c < V; ++c) {
            double v;
            if (c < 3) {
              v = qv[c];
            } else {
              v = 1;
              if (std::abs(x) < 0.025 && std::abs(y) < 0.15 && std::abs(z) < 0.025) v += 10.0;
            }
            h[((static_cast<std::int64_t>(b) * V + c) * e.ent(2) + k) * e.ent(1) * e.ent(0) +
              static_cast<std::int64_t>(j) * e.ent(0) + i] = v;
          }
        }
  }
  return PgenUpload(md, f, h);
}

}  // extern "C"

extern "C" {

pth_status pth_profile_enable(int enable) {
  Prof().on = enable != 0;
  if (!Prof().on) Prof().Clear();
  return PTH_COMPLETE;
}

int pth_profile_report(char *buf, int len) {
  std::map<std::string, std::pair<int, double>> acc;
  for (auto &r : Prof().records) {
    if (hipEventSynchronize(r.stop) != hipSuccess) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) continue;
    auto &a = acc[r.name];
    a.first += 1;
    a.second += ms;
  }
  Prof().Clear();
  std::string out;
  for (auto &kv : acc) {
    char line[256];
    std::snprintf(line, sizeof(line), "%s %d %.6f\n", kv.first.c_str(), kv.second.first,
                  kv.second.second);
    out += line;
  }
  if (buf && len > 0) {
    std::strncpy(buf, out.c_str(), static_cast<size_t>(len) - 1);
    buf[len - 1] = '\0';
  }
  return static_cast<int>(acc.size());
}

}  // extern "C"

// ---- CFL estimate without allocation ---------------------------------------------------
// The workgroup partials of the reduction live in one file-scope device array per GPU,
// `scratch`, instead of the mesh's scratch area. In host code `scratch` names a host-side
// stand-in, never the device memory: the device address of each GPU's copy is taken once,
// with that GPU current, by hipGetSymbolAddress and kept per device; only that address is
// passed to kernels.
namespace {

constexpr int kStaticPartials = 8192;  // 64 KiB per GPU, held until the process ends
__device__ __align__(256) double scratch[kStaticPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

// device address of the current GPU's `scratch` (looked up on the first call per device)
pth_status StaticScratch(double **out) {
  int dev = -1;
  PTH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    SetError("static scratch: device ordinal out of range");
    return PTH_FAIL;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    PTH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PTH_COMPLETE;
}

}  // namespace

extern "C" pth_status pth_burgers_estimate_timestep_noalloc(pth_meshdata *mdh, double cfl,
                                                            double *dt_device) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  const int fu = NeedField(md, "U", 3);
  if (fu < 0) return PTH_FAIL;
  if (!dt_device) {
    SetError("pth_burgers_estimate_timestep_noalloc: dt_device is null");
    return PTH_FAIL;
  }
  const int need = BurgersDtPartials(m->topo);
  if (need > kStaticPartials) {
    SetError("pth_burgers_estimate_timestep_noalloc: the mesh needs " + std::to_string(need) +
             " reduction partials, more than the " + std::to_string(kStaticPartials) +
             " of the static scratch");
    return PTH_FAIL;
  }
  double *partial = nullptr;
  if (StaticScratch(&partial) != PTH_COMPLETE) return PTH_FAIL;
  return BurgersDtToDevice(m->topo, m->fields[fu].ncomp, md->store[fu].data, m->coef, partial,
                           kStaticPartials, cfl, dt_device, m->stream);
}
