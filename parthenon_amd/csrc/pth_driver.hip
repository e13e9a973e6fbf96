// Device-resident Burgers driver: the task order of BurgersDriver::MakeTaskCollection
// (benchmarks/burgers/burgers_driver.cpp:53-148) with the low-storage RK2 integrator
// (src/time_integration/low_storage_integrator.cpp:67-83) and EvolutionDriver's time-step
// control (src/driver/driver.cpp:194-235), with every host decision hoisted out of the stage:
//
//   stage 1: recon+HLL(base) -> [flux correction] -> update base -> "1" -> exchange("1")
//   stage 2: recon+HLL("1")  -> [flux correction] -> update ("1", base) -> base + CFL
//            partials -> exchange(base)
//   clock:   time += dt; dt = min(2 dt, cfl * min)   (all on device)
//
// For cubic blocks of 8 or 16 cells with 4 ghosts (and 4-13 components) the reconstruction
// and HLL run plane by plane in LDS and only the fluxes reach HBM; other shapes keep the face
// states in HBM and evaluate HLL inside the update.
//
// The simulation clock lives in device memory, so a whole cycle has no host synchronization
// and is captured once into a hipGraph and replayed. The exchange inside a cycle uses the same
// persistent pack / unpack / restrict / prolong tables as SendBoundBufs / SetBounds /
// ProlongateBounds.
//
// Containers: the fused update writes only interior cells of the stage's output. Ghost cells of
// that output are refilled by the exchange that follows, and the reference's in-place average
// of "1" in stage 2 is never read again (stage 1 of the next cycle overwrites it), so the
// trajectory of "base" is identical to the task-by-task form.
#include <hip/hip_runtime.h>

#include <cmath>

#include <cfloat>
#include <cstring>
#include <vector>

#include "pth_burgers.hpp"

namespace pth {
namespace {

// The exchange of AddBoundaryExchangeTasks (boundary_communication.cpp:432-445): send + set
// (restriction on multilevel meshes), coarse-buffer boundary conditions, prolongation,
// boundary conditions.
// With rest, the same-level channels of field c.rest_field are left out (the block sweep has
// filled them).
pth_status Exchange(Mesh &m, MeshData &md, hipStream_t s, bool rest = false) {
  if (m.BuildCache(md, PTH_BT_ANY) != PTH_COMPLETE) return PTH_FAIL;
  BoundCache &c = md.cache[PTH_BT_ANY];
  if (c.nbuffers > 0) {
    if (m.topo.multilevel &&
        LaunchRestrictT(c.restrict_send, m.topo, m.geom, s) != PTH_COMPLETE)
      return PTH_FAIL;
    if (rest) {
      if (c.rest_field < 0) {
        SetError("internal: exchange without same-level channels was not built");
        return PTH_FAIL;
      }
      if (LaunchExchangeDirect(c.pack_rest, c.unpack_rest, s) != PTH_COMPLETE) return PTH_FAIL;
    } else if (m.pack_mode == PTH_PACK_PERSISTENT && CanExchangeDirect(c.pack, c.unpack)) {
      // same-level channels by ghost shells where the block shape allows, the rest by regions
      bool shell = false;
      if (m.BuildCrossExchange(md, &shell) != PTH_COMPLETE) return PTH_FAIL;
      if (shell) {
        if (LaunchShell(m.topo, c.shell, m.same_level, c.shell_order, s) != PTH_COMPLETE)
          return PTH_FAIL;
        if (LaunchExchangeDirect(c.pack_cross, c.unpack_cross, s) != PTH_COMPLETE)
          return PTH_FAIL;
      } else if (LaunchExchangeDirect(c.pack, c.unpack, s) != PTH_COMPLETE) {
        return PTH_FAIL;
      }
    } else {
      if (LaunchPack(c.pack, nullptr, s, true, m.pack_mode) != PTH_COMPLETE) return PTH_FAIL;
      if (LaunchPack(c.unpack, nullptr, s, false, m.pack_mode) != PTH_COMPLETE) return PTH_FAIL;
    }
    if (m.topo.multilevel &&
        LaunchRestrictT(c.restrict_set, m.topo, m.geom, s) != PTH_COMPLETE)
      return PTH_FAIL;
  }
  if (m.topo.multilevel) {
    if (LaunchBoundaryConditions(c.bc[1], s) != PTH_COMPLETE) return PTH_FAIL;
    if (c.nbuffers > 0 && LaunchProlongT(c.prolong, m.topo, m.geom, s, true) != PTH_COMPLETE)
      return PTH_FAIL;
    if (c.nbuffers > 0 &&
        LaunchProlongT(c.prolong_linear, m.topo, m.geom, s, false) != PTH_COMPLETE)
      return PTH_FAIL;
  }
  return LaunchBoundaryConditions(c.bc[0], s);
}

struct Ctx {
  Mesh *m;
  MeshData *base, *one;
  int fu, fd, V;
  const FluxCorrection *fc = nullptr;  // multilevel: flux-correction patches of U
  const int *push = nullptr;           // sweep: same-level neighbours it writes ghosts of
  bool split = false;                  // sweep: split stages (Mesh::sweep_split)
};

// Lists A and B of the split sweep: A holds the blocks whose every neighbour is at their own
// level and that touch no physical face (no cross-level channel, flux correction or boundary
// condition writes or reads them), B the others, each in block order.
pth_status SplitLists(Mesh &m) {
  if (m.split_gen == m.generation) return PTH_COMPLETE;
  const Topology &t = m.topo;
  std::vector<int> a, b;
  for (int q = 0; q < static_cast<int>(t.blocks.size()); ++q) {
    const Block &B = t.blocks[q];
    bool inner = true;
    for (const Link &l : B.links) inner = inner && l.level == B.at.level;
    for (int d = 0; d < 3; ++d) {
      if (!t.active[d] || t.periodic[d]) continue;
      const double tol = 1e-12 * (t.xmax[d] - t.xmin[d]);
      inner = inner && std::fabs(B.xmin[d] - t.xmin[d]) > tol &&
              std::fabs(B.xmax[d] - t.xmax[d]) > tol;
    }
    (inner ? a : b).push_back(q);
  }
  if (m.split_list) (void)hipFree(m.split_list);
  m.split_list = nullptr;
  m.split_nA = static_cast<int>(a.size());
  m.split_nB = static_cast<int>(b.size());
  a.insert(a.end(), b.begin(), b.end());
  if (!a.empty()) {
    PTH_HIP(hipMalloc(&m.split_list, sizeof(int) * a.size()));
    PTH_HIP(hipMemcpy(m.split_list, a.data(), sizeof(int) * a.size(), hipMemcpyHostToDevice));
  }
  m.split_gen = m.generation;
  return PTH_COMPLETE;
}

pth_status Setup(Mesh *m, Ctx &x) {
  if (!m->finalized) {
    SetError("mesh is not finalized");
    return PTH_FAIL;
  }
  x.m = m;
  x.fu = m->FieldIndex("U");
  x.fd = m->FieldIndex("derived");
  if (x.fu < 0 || m->fields[x.fu].ncomp < 4) {
    SetError("Burgers driver needs field U with >= 4 components");
    return PTH_FAIL;
  }
  x.V = m->fields[x.fu].ncomp;
  if (m->topo.multilevel) {
    if (!(m->fields[x.fu].flags & PTH_WITH_FLUXES)) {
      SetError("multilevel Burgers needs U declared WithFluxes (flux correction)");
      return PTH_FAIL;
    }
    x.fc = m->FluxPatches(x.fu);
    if (!x.fc) return PTH_FAIL;
  }
  x.base = m->Get("base");
  x.one = m->Get("1");
  if (!x.base || !x.one) return PTH_FAIL;
  if (!m->run_stream) PTH_HIP(hipStreamCreateWithFlags(&m->run_stream, hipStreamNonBlocking));
  if (!m->side_stream)
    PTH_HIP(hipStreamCreateWithFlags(&m->side_stream, hipStreamNonBlocking));
  for (hipEvent_t &e : m->fork_ev)
    if (!e) PTH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (!m->dtvar) PTH_HIP(hipMalloc(&m->dtvar, sizeof(double) * kClockSlots));
  const std::int64_t nd = x.base->store[x.fu].ndata;
  const bool planes = x.base->store[x.fu].flux && ReconFluxSupported(m->topo, x.V);
  const bool sweep = m->run_sweep && BurgersSweepSupported(m->topo, x.V);
  if (!planes && !sweep && (!m->face || m->face_V != x.V)) {  // face states of the other path
    if (m->face) (void)hipFree(m->face);
    m->face = nullptr;
    PTH_HIP(hipMalloc(&m->face, sizeof(double) * 6 * nd));
    PTH_HIP(hipMemset(m->face, 0, sizeof(double) * 6 * nd));
    m->face_V = x.V;
  }
  if (FusedUpdatePartials(m->topo) > m->scratch_n) {
    SetError("too many CFL partials for the scratch area");
    return PTH_FAIL;
  }
  // the sweep writes U's same-level ghost cells itself when the rest of the exchange can run
  // without those channels (host tables, built before any graph capture)
  x.push = nullptr;
  if (sweep) {
    bool ok_base = false, ok_one = false;
    if (m->BuildRestExchange(*x.base, x.fu, &ok_base) != PTH_COMPLETE ||
        m->BuildRestExchange(*x.one, x.fu, &ok_one) != PTH_COMPLETE ||
        m->SameLevelNeighbors() != PTH_COMPLETE)
      return PTH_FAIL;
    if (ok_base && ok_one && m->pack_mode == PTH_PACK_PERSISTENT) x.push = m->same_level;
    if (x.push && m->sweep_split && m->topo.multilevel && m->topo.fine.n[0] == 16 &&
        m->topo.fine.g[0] == 4) {
      if (SplitLists(*m) != PTH_COMPLETE) return PTH_FAIL;
      x.split = m->split_nA > 0 && m->split_nB > 0;
    }
  }
  return PTH_COMPLETE;
}

pth_status Stage(Ctx &x, int stage, hipStream_t s) {
  Mesh &m = *x.m;
  MeshData &mc0 = stage == 1 ? *x.base : *x.one;
  MeshData &mc1 = stage == 1 ? *x.one : *x.base;
  const double beta = stage == 1 ? 1.0 : 0.5;
  double *derived = x.fd >= 0 ? x.base->store[x.fd].data : nullptr;
  double *flux = x.base->store[x.fu].flux;  // scratch of both stages
  if (flux && ReconFluxSupported(m.topo, x.V)) {
    // plane-tiled reconstruction + HLL into the flux array, the flux-correction tasks on the
    // stored fluxes (LoadAndSendFluxCorrections, SetFluxCorrections), then the update
    if (BurgersReconFlux(m.topo, x.V, mc0.store[x.fu].data, flux, m.run_recon, 0, m.rf_ppw[0], s) !=
        PTH_COMPLETE)
      return PTH_FAIL;
    if (x.fc && (FlxcorSend(m.topo, *x.fc, flux, s) != PTH_COMPLETE ||
                 FlxcorSet(m.topo, *x.fc, flux, s) != PTH_COMPLETE))
      return PTH_FAIL;
    if (BurgersFluxUpdate(m.topo, x.V, flux, mc0.store[x.fu].data, x.base->store[x.fu].data,
                          mc1.store[x.fu].data, derived, stage == 2 ? m.scratch : nullptr,
                          m.coef, beta, m.dtvar, s) != PTH_COMPLETE)
      return PTH_FAIL;
  } else {
    if (BurgersReconTasks(m.topo, x.V, mc0.store[x.fu].data, m.face, m.run_recon, s) !=
        PTH_COMPLETE)
      return PTH_FAIL;
    // LoadAndSendFluxCorrections / SetFluxCorrections: the restricted fine fluxes land in
    // the correction store that the fused update reads in place of its coarse-face fluxes
    if (x.fc && FlxcorFromFaces(m.topo, x.V, *x.fc, m.face, s) != PTH_COMPLETE)
      return PTH_FAIL;
    if (BurgersFusedUpdate(m.topo, x.V, m.face, mc0.store[x.fu].data,
                           x.base->store[x.fu].data, mc1.store[x.fu].data, derived, nullptr,
                           stage == 2 ? m.scratch : nullptr, m.coef, beta, m.dtvar, x.fc,
                           s) != PTH_COMPLETE)
      return PTH_FAIL;
  }
  if (Exchange(m, mc1, s) != PTH_COMPLETE) return PTH_FAIL;
  if (stage == 2)
    return BurgersClock(m.scratch, FusedUpdatePartials(m.topo), m.run_cfl, m.dtvar, 1, m.run_lim,
                        s);
  return PTH_COMPLETE;
}

bool Swept(const Ctx &x) {
  return x.m->run_sweep && BurgersSweepSupported(x.m->topo, x.V);
}

bool Overlapped(const Ctx &x) {
  return !Swept(x) && x.m->run_overlap && x.base->store[x.fu].flux &&
         ReconFluxSupported(x.m->topo, x.V);
}

// The stage as one block-sweep launch (after the flux-correction send on multilevel meshes,
// whose restricted fine fluxes the sweep reads for the faces that have a finer neighbour),
// then the exchange of its output.
pth_status StageSweep(Ctx &x, int stage, hipStream_t s) {
  Mesh &m = *x.m;
  MeshData &mc0 = stage == 1 ? *x.base : *x.one;
  MeshData &mc1 = stage == 1 ? *x.one : *x.base;
  const double beta = stage == 1 ? 1.0 : 0.5;
  double *derived = x.fd >= 0 ? x.base->store[x.fd].data : nullptr;
  const double *U = mc0.store[x.fu].data;
  if (x.split) {
    // side stream: the cross-level part of the exchange of this stage's input (deferred from
    // the end of the previous stage) and the flux-correction send, beside the sweep of list
    // A, whose blocks they neither read nor write; then the sweep of list B, which needs them
    PTH_HIP(hipEventRecord(m.fork_ev[0], s));
    PTH_HIP(hipStreamWaitEvent(m.side_stream, m.fork_ev[0], 0));
    if (Exchange(m, mc0, m.side_stream, true) != PTH_COMPLETE) return PTH_FAIL;
    // the flux-correction send reads the interface blocks' input with its ghosts and writes
    // the correction store, which only sweep B reads
    if (x.fc && FlxcorDirect(m.topo, x.V, *x.fc, U, m.run_recon, m.side_stream) != PTH_COMPLETE)
      return PTH_FAIL;
    PTH_HIP(hipEventRecord(m.fork_ev[1], m.side_stream));
    int ga = 0, gb = 0;
    double *part = stage == 2 ? m.scratch : nullptr;
    if (BurgersSweep(m.topo, x.V, U, x.base->store[x.fu].data, mc1.store[x.fu].data, derived,
                     part, static_cast<int>(m.scratch_n) / 2, m.coef, beta, m.dtvar,
                     m.run_recon, x.fc, m.sweep_waves, m.sweep_spread, x.push, &ga, s,
                     m.split_list, m.split_nA) != PTH_COMPLETE)
      return PTH_FAIL;
    PTH_HIP(hipStreamWaitEvent(s, m.fork_ev[1], 0));
    if (BurgersSweep(m.topo, x.V, U, x.base->store[x.fu].data, mc1.store[x.fu].data, derived,
                     part ? part + ga : nullptr, static_cast<int>(m.scratch_n) / 2, m.coef,
                     beta, m.dtvar, m.run_recon, x.fc, m.sweep_waves, m.sweep_spread, x.push,
                     &gb, s, m.split_list + m.split_nA, m.split_nB) != PTH_COMPLETE)
      return PTH_FAIL;
    if (stage == 2) return BurgersClock(m.scratch, ga + gb, m.run_cfl, m.dtvar, 1, m.run_lim, s);
    return PTH_COMPLETE;
  }
  if (x.fc && FlxcorDirect(m.topo, x.V, *x.fc, U, m.run_recon, s) != PTH_COMPLETE)
    return PTH_FAIL;
  int groups = 0;
  if (BurgersSweep(m.topo, x.V, U, x.base->store[x.fu].data, mc1.store[x.fu].data, derived,
                   stage == 2 ? m.scratch : nullptr, static_cast<int>(m.scratch_n), m.coef, beta,
                   m.dtvar, m.run_recon, x.fc, m.sweep_waves, m.sweep_spread, x.push, &groups,
                   s) !=
      PTH_COMPLETE)
    return PTH_FAIL;
  if (Exchange(m, mc1, s, x.push != nullptr) != PTH_COMPLETE) return PTH_FAIL;
  if (stage == 2) return BurgersClock(m.scratch, groups, m.run_cfl, m.dtvar, 1, m.run_lim, s);
  return PTH_COMPLETE;
}

// The stage with its input's boundary exchange moved from the end of the previous stage to
// the start of this one, beside the reconstruction of the cells whose stencils read no ghost:
//   main stream: recon+HLL (ghost-free faces)  |  side stream: exchange(input)
//   join -> recon+HLL (remaining faces) -> [flux correction] -> update -> [clock]
// The two branches touch disjoint data: the exchange reads interior cells and writes ghost
// cells and coarse buffers, the first reconstruction reads interior cells and writes fluxes.
pth_status StageOverlap(Ctx &x, int stage, hipStream_t s) {
  Mesh &m = *x.m;
  MeshData &mc0 = stage == 1 ? *x.base : *x.one;
  MeshData &mc1 = stage == 1 ? *x.one : *x.base;
  const double beta = stage == 1 ? 1.0 : 0.5;
  double *derived = x.fd >= 0 ? x.base->store[x.fd].data : nullptr;
  double *flux = x.base->store[x.fu].flux;
  const double *U = mc0.store[x.fu].data;
  PTH_HIP(hipEventRecord(m.fork_ev[0], s));
  PTH_HIP(hipStreamWaitEvent(m.side_stream, m.fork_ev[0], 0));
  if (Exchange(m, mc0, m.side_stream) != PTH_COMPLETE) return PTH_FAIL;
  if (BurgersReconFlux(m.topo, x.V, U, flux, m.run_recon, 1, m.rf_ppw[1], s) != PTH_COMPLETE)
    return PTH_FAIL;
  PTH_HIP(hipEventRecord(m.fork_ev[1], m.side_stream));
  PTH_HIP(hipStreamWaitEvent(s, m.fork_ev[1], 0));
  if (BurgersReconFlux(m.topo, x.V, U, flux, m.run_recon, 2, m.rf_ppw[2], s) != PTH_COMPLETE)
    return PTH_FAIL;
  if (x.fc && (FlxcorSend(m.topo, *x.fc, flux, s) != PTH_COMPLETE ||
               FlxcorSet(m.topo, *x.fc, flux, s) != PTH_COMPLETE))
    return PTH_FAIL;
  if (BurgersFluxUpdate(m.topo, x.V, flux, U, x.base->store[x.fu].data, mc1.store[x.fu].data,
                        derived, stage == 2 ? m.scratch : nullptr, m.coef, beta, m.dtvar,
                        s) != PTH_COMPLETE)
    return PTH_FAIL;
  if (stage == 2)
    return BurgersClock(m.scratch, FusedUpdatePartials(m.topo), m.run_cfl, m.dtvar, 1, m.run_lim,
                        s);
  return PTH_COMPLETE;
}

// Picks the planes per workgroup of each reconstruction part by timing every choice on the
// mesh's own data (the flux array is scratch between stages, so the trial launches change
// nothing); the choice changes the work split only, never a result.
pth_status TuneReconFlux(Ctx &x, hipStream_t s) {
  Mesh &m = *x.m;
  double *flux = x.base->store[x.fu].flux;
  const bool planes = flux && ReconFluxSupported(m.topo, x.V);
  const double *U = x.base->store[x.fu].data;
  hipEvent_t ev[2];
  PTH_HIP(hipEventCreate(&ev[0]));
  PTH_HIP(hipEventCreate(&ev[1]));
  const bool prof = Prof().on;
  Prof().on = false;
  pth_status st = PTH_COMPLETE;
  for (int part = 0; planes && part < 3 && st == PTH_COMPLETE; ++part) {
    float best = 0.0f;
    for (int ppw : {1, 2, 4}) {
      st = BurgersReconFlux(m.topo, x.V, U, flux, m.run_recon, part, ppw, s);  // warm-up
      if (st != PTH_COMPLETE) break;
      if (hipEventRecord(ev[0], s) != hipSuccess) st = PTH_FAIL;
      for (int r = 0; r < 3 && st == PTH_COMPLETE; ++r)
        st = BurgersReconFlux(m.topo, x.V, U, flux, m.run_recon, part, ppw, s);
      float ms = 0.0f;
      if (st != PTH_COMPLETE || hipEventRecord(ev[1], s) != hipSuccess ||
          hipEventSynchronize(ev[1]) != hipSuccess ||
          hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
        if (st == PTH_COMPLETE) SetError("reconstruction tuning: event timing failed");
        st = PTH_FAIL;
        break;
      }
      if (ppw == 1 || ms < best) {
        best = ms;
        m.rf_ppw[part] = ppw;
      }
    }
  }
  // block sweep: the register budget (built for 3 or 4 waves per SIMD); the trial writes only
  // the interior of stage "1", which the next stage overwrites
  if (st == PTH_COMPLETE && BurgersSweepSupported(m.topo, x.V)) {
    float best = 0.0f;
    // the register-ring kernel of 16^3 blocks is built for 3 or 2 waves per SIMD, the general
    // one for 4 or 3
    const bool ring = m.topo.fine.n[0] == 16 && m.topo.fine.g[0] == 4;
    const int budgets[2] = {ring ? 3 : 4, ring ? 2 : 3};
    for (int waves : budgets) {
      st = BurgersSweep(m.topo, x.V, U, U, x.one->store[x.fu].data, nullptr, nullptr,
                        static_cast<int>(m.scratch_n), m.coef, 1.0, m.dtvar, m.run_recon, x.fc,
                        waves, 1, nullptr, nullptr, s);
      if (st != PTH_COMPLETE) break;
      if (hipEventRecord(ev[0], s) != hipSuccess) st = PTH_FAIL;
      for (int r = 0; r < 3 && st == PTH_COMPLETE; ++r)
        st = BurgersSweep(m.topo, x.V, U, U, x.one->store[x.fu].data, nullptr, nullptr,
                          static_cast<int>(m.scratch_n), m.coef, 1.0, m.dtvar, m.run_recon,
                          x.fc, waves, 1, nullptr, nullptr, s);
      float ms = 0.0f;
      if (st != PTH_COMPLETE || hipEventRecord(ev[1], s) != hipSuccess ||
          hipEventSynchronize(ev[1]) != hipSuccess ||
          hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
        if (st == PTH_COMPLETE) SetError("sweep tuning: event timing failed");
        st = PTH_FAIL;
        break;
      }
      if (waves == budgets[0] || ms < best) {
        best = ms;
        m.sweep_waves = waves;
      }
    }
    // then how many resident workgroups share the planes (longer segments on small meshes)
    m.sweep_spread = 1;
    for (int spread : {1, 2, 3}) {
      if (st != PTH_COMPLETE) break;
      for (int r = 0; r < 4 && st == PTH_COMPLETE; ++r) {
        if (r == 1 && hipEventRecord(ev[0], s) != hipSuccess) st = PTH_FAIL;
        if (st == PTH_COMPLETE)
          st = BurgersSweep(m.topo, x.V, U, U, x.one->store[x.fu].data, nullptr, nullptr,
                            static_cast<int>(m.scratch_n), m.coef, 1.0, m.dtvar, m.run_recon,
                            x.fc, m.sweep_waves, spread, nullptr, nullptr, s);
      }
      float ms = 0.0f;
      if (st != PTH_COMPLETE || hipEventRecord(ev[1], s) != hipSuccess ||
          hipEventSynchronize(ev[1]) != hipSuccess ||
          hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
        if (st == PTH_COMPLETE) SetError("sweep tuning: event timing failed");
        st = PTH_FAIL;
        break;
      }
      if (spread == 1 || ms < best) {
        best = ms;
        m.sweep_spread = spread;
      }
    }
  }
  Prof().on = prof;
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (st == PTH_COMPLETE) m.rf_tuned = true;
  return st;
}

pth_status Cycle(Ctx &x, hipStream_t s);

// Picks the stage form (block sweep; plane reconstruction with the exchange on a second
// stream; plane reconstruction then update then exchange) by timing whole cycles of each on
// the mesh's own state, which is saved first and restored afterwards (U of "base", the
// derived field and the device clock). The two plane forms give the same bits; the sweep is
// built with FMA contraction and agrees with them to rounding (about 1e-15 relative), so
// with form_exact set only the plane forms are candidates and the result cannot depend on
// which form the timing picked.
pth_status TuneForm(Ctx &x, hipStream_t s) {
  Mesh &m = *x.m;
  struct Form {
    int sweep, overlap;
  };
  std::vector<Form> forms;
  if (BurgersSweepSupported(m.topo, x.V) &&
      !(m.form_exact && x.base->store[x.fu].flux && ReconFluxSupported(m.topo, x.V)))
    forms.push_back({1, 0});
  if (x.base->store[x.fu].flux && ReconFluxSupported(m.topo, x.V)) {
    forms.push_back({0, 1});
    forms.push_back({0, 0});
  }
  if (forms.size() < 2) {
    m.form_tuned = true;
    return PTH_COMPLETE;
  }
  const std::int64_t nd = x.base->store[x.fu].ndata;
  double *U = x.base->store[x.fu].data;
  double *D = x.fd >= 0 ? x.base->store[x.fd].data : nullptr;
  const std::int64_t ndd = x.fd >= 0 ? x.base->store[x.fd].ndata : 0;
  double *save = nullptr;
  PTH_HIP(hipMalloc(&save, sizeof(double) * (nd + ndd + kClockSlots)));
  hipEvent_t ev[2] = {nullptr, nullptr};
  pth_status st = PTH_COMPLETE;
  auto copy = [&](double *dst, const double *src, std::int64_t n) {
    if (st == PTH_COMPLETE && n > 0 &&
        hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      SetError("stage-form tuning: state copy failed");
      st = PTH_FAIL;
    }
  };
  copy(save, U, nd);
  if (D) copy(save + nd, D, ndd);
  copy(save + nd + ndd, m.dtvar, kClockSlots);
  if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
    SetError("stage-form tuning: event creation failed");
    st = PTH_FAIL;
  }
  const bool prof = Prof().on;
  Prof().on = false;
  const int keep_sweep = m.run_sweep, keep_overlap = m.run_overlap;
  float best = 0.0f;
  Form pick = forms[0];
  for (const Form &f : forms) {
    if (st != PTH_COMPLETE) break;
    m.run_sweep = f.sweep;
    m.run_overlap = f.overlap;
    copy(U, save, nd);
    copy(m.dtvar, save + nd + ndd, kClockSlots);
    if (st == PTH_COMPLETE) st = Cycle(x, s);  // warm-up
    if (st == PTH_COMPLETE && hipEventRecord(ev[0], s) != hipSuccess) st = PTH_FAIL;
    for (int c = 0; c < 2 && st == PTH_COMPLETE; ++c) st = Cycle(x, s);
    float ms = 0.0f;
    if (st != PTH_COMPLETE || hipEventRecord(ev[1], s) != hipSuccess ||
        hipEventSynchronize(ev[1]) != hipSuccess ||
        hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) {
      if (st == PTH_COMPLETE) SetError("stage-form tuning: event timing failed");
      st = PTH_FAIL;
      break;
    }
    if (&f == &forms[0] || ms < best) {
      best = ms;
      pick = f;
    }
  }
  Prof().on = prof;
  copy(U, save, nd);
  if (D) copy(D, save + nd, ndd);
  copy(m.dtvar, save + nd + ndd, kClockSlots);
  if (st == PTH_COMPLETE && hipStreamSynchronize(s) != hipSuccess) {
    SetError("stage-form tuning: synchronize failed");
    st = PTH_FAIL;
  }
  (void)hipFree(save);
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  if (st != PTH_COMPLETE) {
    m.run_sweep = keep_sweep;
    m.run_overlap = keep_overlap;
    return st;
  }
  m.run_sweep = pick.sweep;
  m.run_overlap = pick.overlap;
  m.form_tuned = true;
  return PTH_COMPLETE;
}

pth_status Cycle(Ctx &x, hipStream_t s) {
  if (Swept(x)) {
    if (StageSweep(x, 1, s) != PTH_COMPLETE) return PTH_FAIL;
    return StageSweep(x, 2, s);
  }
  if (Overlapped(x)) {
    if (StageOverlap(x, 1, s) != PTH_COMPLETE) return PTH_FAIL;
    return StageOverlap(x, 2, s);
  }
  if (Stage(x, 1, s) != PTH_COMPLETE) return PTH_FAIL;
  return Stage(x, 2, s);
}

// ncycles eager cycles, or replays of the cycle captured on first use
pth_status RunCycles(Mesh &m, Ctx &x, hipStream_t s, int ncycles, bool use_graph) {
  if (!use_graph) {
    for (int c = 0; c < ncycles; ++c)
      if (Cycle(x, s) != PTH_COMPLETE) return PTH_FAIL;
    return PTH_COMPLETE;
  }
  if (!m.graph_exec) {
    // tables must exist before capture (their uploads are synchronous copies)
    if (m.BuildCache(*x.base, PTH_BT_ANY) != PTH_COMPLETE) return PTH_FAIL;
    if (m.BuildCache(*x.one, PTH_BT_ANY) != PTH_COMPLETE) return PTH_FAIL;
    bool shell = false;
    if (m.BuildCrossExchange(*x.base, &shell) != PTH_COMPLETE) return PTH_FAIL;
    if (m.BuildCrossExchange(*x.one, &shell) != PTH_COMPLETE) return PTH_FAIL;
    const bool prof = Prof().on;
    Prof().on = false;  // no event records inside the captured cycle
    PTH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const pth_status st = Cycle(x, s);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    Prof().on = prof;
    if (st != PTH_COMPLETE) {
      if (g) (void)hipGraphDestroy(g);
      return PTH_FAIL;
    }
    PTH_HIP(e);
    m.graph = g;
    PTH_HIP(hipGraphInstantiate(&m.graph_exec, g, nullptr, nullptr, 0));
  }
  {
    ProfScope prof("graph_cycle", s);
    for (int c = 0; c < ncycles; ++c) PTH_HIP(hipGraphLaunch(m.graph_exec, s));
  }
  return PTH_COMPLETE;
}

}  // namespace
}  // namespace pth

using namespace pth;

extern "C" {

pth_status pth_exchange_fused(pth_meshdata *mdh, int bt) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (bt == PTH_BT_LOCAL) bt = PTH_BT_ANY;
  if (bt == PTH_BT_NONLOCAL) return PTH_COMPLETE;
  if (bt != PTH_BT_ANY) {
    SetError("fused exchange serves the cell-variable bound types any/local/nonlocal");
    return PTH_FAIL;
  }
  if (!m->finalized) {
    SetError("mesh is not finalized");
    return PTH_FAIL;
  }
  return Exchange(*m, *md, m->stream);
}

pth_status pth_burgers_init(pth_mesh *mh, double cfl, int recon) {
  Mesh *m = Impl(mh);
  Ctx x;
  if (Setup(m, x) != PTH_COMPLETE) return PTH_FAIL;
  for (int d = 0; d < m->topo.ndim; ++d)
    if (m->topo.fine.g[d] < (recon == 0 ? 4 : 2)) {
      SetError("weno5 reconstruction requires 4 or more ghost cells");
      return PTH_FAIL;
    }
  m->run_cfl = cfl;
  m->run_recon = recon;
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  hipStream_t s = m->run_stream;
  // order after any work the caller queued on the mesh stream
  PTH_HIP(hipStreamSynchronize(m->stream));
  const double init[kClockSlots] = {DBL_MAX, 0.0, 0.0, DBL_MAX, DBL_MAX};
  PTH_HIP(hipMemcpyAsync(m->dtvar, init, sizeof(init), hipMemcpyHostToDevice, s));
  // Mesh::Initialize after the problem generator: exchange, FillDerived, first time step
  if (Exchange(*m, *x.base, s) != PTH_COMPLETE) return PTH_FAIL;
  if (x.fd >= 0 && BurgersDerived(m->topo, x.V, x.base->store[x.fu].data,
                                  x.base->store[x.fd].data, s) != PTH_COMPLETE)
    return PTH_FAIL;
  const int np = static_cast<int>(m->scratch_n);
  if (BurgersDtMin(m->topo, x.V, x.base->store[x.fu].data, m->coef, m->scratch, np,
                   m->scratch + m->scratch_n, s) != PTH_COMPLETE)
    return PTH_FAIL;
  // the estimate of BurgersDtMin is the un-scaled minimum; apply the CFL and the clock rule
  if (BurgersClock(m->scratch + m->scratch_n, 1, cfl, m->dtvar, 0, m->run_lim, s) !=
      PTH_COMPLETE)
    return PTH_FAIL;
  if (!m->rf_tuned && TuneReconFlux(x, s) != PTH_COMPLETE) return PTH_FAIL;
  if (m->form_auto && !m->form_tuned && TuneForm(x, s) != PTH_COMPLETE) return PTH_FAIL;
  m->tuned_blocks = static_cast<int>(m->topo.blocks.size());
  PTH_HIP(hipStreamSynchronize(s));
  m->run_ready = true;
  return PTH_COMPLETE;
}

pth_status pth_burgers_form(pth_mesh *mh, int *sweep, int *overlap) {
  Mesh *m = Impl(mh);
  if (sweep) *sweep = m->run_sweep;
  if (overlap) *overlap = m->run_overlap;
  return m->form_auto && !m->form_tuned ? PTH_INCOMPLETE : PTH_COMPLETE;
}

pth_status pth_burgers_tuning(pth_mesh *mh, int planes[4]) {
  Mesh *m = Impl(mh);
  for (int p = 0; p < 3; ++p) planes[p] = m->rf_ppw[p];
  planes[3] = m->sweep_waves;
  return m->rf_tuned ? PTH_COMPLETE : PTH_INCOMPLETE;
}

pth_status pth_burgers_set_tuning(pth_mesh *mh, const int planes[4]) {
  Mesh *m = Impl(mh);
  for (int p = 0; p < 3; ++p)
    if (planes[p] != 1 && planes[p] != 2 && planes[p] != 4) {
      SetError("planes per workgroup must be 1, 2 or 4");
      return PTH_FAIL;
    }
  if (planes[3] < 2 || planes[3] > 4) {
    SetError("block-sweep waves per SIMD must be 2, 3 or 4");
    return PTH_FAIL;
  }
  for (int p = 0; p < 3; ++p) m->rf_ppw[p] = planes[p];
  m->sweep_waves = planes[3];
  m->rf_tuned = true;  // a caller's choice is not re-tuned
  m->rf_auto = false;
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  return PTH_COMPLETE;
}

pth_status pth_burgers_run_cycles(pth_mesh *mh, int ncycles, int use_graph) {
  Mesh *m = Impl(mh);
  if (!m->run_ready) {
    SetError("call pth_burgers_init first");
    return PTH_FAIL;
  }
  Ctx x;
  if (Setup(m, x) != PTH_COMPLETE) return PTH_FAIL;
  hipStream_t s = m->run_stream;
  for (hipEvent_t &e : m->order_ev)
    if (!e) PTH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // the cycles start after work already queued on the mesh stream (uploads, task calls), and
  // work queued on the mesh stream afterwards (downloads, task calls) starts after the cycles
  PTH_HIP(hipEventRecord(m->order_ev[0], m->stream));
  PTH_HIP(hipStreamWaitEvent(s, m->order_ev[0], 0));
  if (RunCycles(*m, x, s, ncycles, use_graph != 0) != PTH_COMPLETE) return PTH_FAIL;
  // with the exchange at the start of each stage, the last output's ghosts are still stale
  if (ncycles > 0 && Overlapped(x) && Exchange(*m, *x.base, s) != PTH_COMPLETE) return PTH_FAIL;
  // the split sweep defers the cross-level part of each stage's exchange to the next stage
  if (ncycles > 0 && Swept(x) && x.split && Exchange(*m, *x.base, s, true) != PTH_COMPLETE)
    return PTH_FAIL;
  PTH_HIP(hipEventRecord(m->order_ev[1], s));
  PTH_HIP(hipStreamWaitEvent(m->stream, m->order_ev[1], 0));
  return PTH_COMPLETE;
}

pth_status pth_burgers_set_sweep(pth_mesh *mh, int on) {
  Mesh *m = Impl(mh);
  m->run_sweep = on != 0;
  m->form_auto = false;
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  return PTH_COMPLETE;
}

pth_status pth_burgers_set_sweep_split(pth_mesh *mh, int on) {
  Mesh *m = Impl(mh);
  m->sweep_split = on != 0;
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  return PTH_COMPLETE;
}

pth_status pth_burgers_set_bit_exact(pth_mesh *mh, int on) {
  Mesh *m = Impl(mh);
  if (m->run_ready && m->form_auto) {
    SetError("pth_burgers_set_bit_exact must precede pth_burgers_init");
    return PTH_FAIL;
  }
  m->form_exact = on != 0;
  return PTH_COMPLETE;
}

pth_status pth_burgers_set_overlap(pth_mesh *mh, int on) {
  Mesh *m = Impl(mh);
  m->run_overlap = on != 0;
  m->form_auto = false;
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  return PTH_COMPLETE;
}

pth_status pth_burgers_after_remesh(pth_mesh *mh) {
  Mesh *m = Impl(mh);
  if (!m->run_ready) {
    SetError("call pth_burgers_init first");
    return PTH_FAIL;
  }
  Ctx x;
  if (Setup(m, x) != PTH_COMPLETE) return PTH_FAIL;
  hipStream_t s = m->run_stream;
  PTH_HIP(hipStreamSynchronize(m->stream));
  // FillDerived, then InitializeBlockTimeSteps and SetGlobalTimeStep from the last cycle's dt
  if (x.fd >= 0 && BurgersDerived(m->topo, x.V, x.base->store[x.fu].data,
                                  x.base->store[x.fd].data, s) != PTH_COMPLETE)
    return PTH_FAIL;
  const int np = static_cast<int>(m->scratch_n);
  if (BurgersDtMin(m->topo, x.V, x.base->store[x.fu].data, m->coef, m->scratch, np,
                   m->scratch + m->scratch_n, s) != PTH_COMPLETE)
    return PTH_FAIL;
  if (BurgersClock(m->scratch + m->scratch_n, 1, m->run_cfl, m->dtvar, 2, m->run_lim, s) !=
      PTH_COMPLETE)
    return PTH_FAIL;
  // choices timed on a block count at least twice smaller or larger are timed again (the
  // tuners restore U, the derived field and the clock)
  const int nb = static_cast<int>(m->topo.blocks.size());
  if ((m->rf_auto || m->form_auto) && (nb >= 2 * m->tuned_blocks || 2 * nb <= m->tuned_blocks)) {
    if (m->rf_auto) m->rf_tuned = false;
    if (m->form_auto) {
      m->form_tuned = false;
      m->run_sweep = 1;  // so that Setup builds the sweep's tables for its trial
    }
    Ctx y;
    if (Setup(m, y) != PTH_COMPLETE) return PTH_FAIL;
    if (!m->rf_tuned && TuneReconFlux(y, s) != PTH_COMPLETE) return PTH_FAIL;
    if (m->form_auto && !m->form_tuned && TuneForm(y, s) != PTH_COMPLETE) return PTH_FAIL;
    m->tuned_blocks = nb;
    if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->graph_exec = nullptr;
    m->graph = nullptr;
  }
  PTH_HIP(hipStreamSynchronize(s));
  return PTH_COMPLETE;
}

pth_status pth_burgers_clock(pth_mesh *mh, double *dt, double *time) {
  Mesh *m = Impl(mh);
  if (!m->run_ready) {
    SetError("call pth_burgers_init first");
    return PTH_FAIL;
  }
  double v[kClockSlots];
  PTH_HIP(hipStreamSynchronize(m->run_stream));
  PTH_HIP(hipMemcpy(v, m->dtvar, sizeof(v), hipMemcpyDeviceToHost));
  if (dt) *dt = v[4];
  if (time) *time = v[1];
  return PTH_COMPLETE;
}

pth_status pth_burgers_set_time_limits(pth_mesh *mh, const double lim[6], int dt_init_force) {
  Mesh *m = Impl(mh);
  if (!lim) {
    SetError("time limits: null array");
    return PTH_FAIL;
  }
  ClockLimits L;
  L.tlim = lim[0];
  L.dt_init = lim[1];
  L.dt_user = lim[2];
  L.dt_force = lim[3];
  L.dt_floor = lim[4];
  L.dt_ceil = lim[5];
  L.dt_init_force = dt_init_force != 0;
  m->run_lim = L;
  // the limits travel in the clock launches' arguments: a captured cycle is captured again
  if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
  if (m->graph) (void)hipGraphDestroy(m->graph);
  m->graph_exec = nullptr;
  m->graph = nullptr;
  return PTH_COMPLETE;
}

pth_status pth_burgers_synchronize(pth_mesh *mh) {
  Mesh *m = Impl(mh);
  if (m->run_stream) PTH_HIP(hipStreamSynchronize(m->run_stream));
  return PTH_COMPLETE;
}

}  // extern "C"
