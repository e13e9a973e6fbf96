// A C++ driver written against the reference's task-level names, run through
// include/parthenon_amd.hpp: the task order of BurgersDriver::MakeTaskCollection
// (benchmarks/burgers/burgers_driver.cpp:53-148) on the header's TaskList, the RK2
// integrator (low_storage_integrator.cpp:67-83) and the driver's dt rule (driver.cpp:210-235).
//
// usage: burgers_task_driver <ncycles> <out.bin> [sphere_radius] [open|periodic] [adaptive]
// writes the final "base" U (float64, [block][comp][k][j][i]) followed by dt and time. With a
// sphere radius the mesh is statically refined and the stage adds the flux-correction tasks
// (burgers_driver.cpp:89-96) and ProlongateBounds after the sets (the library exchange order
// of AddBoundaryExchangeTasks, boundary_communication.cpp:406-445). "open" makes x1 and x3
// non-periodic (inner x1 reflecting, the other three faces outflow) and adds the boundary
// condition tasks (burgers_driver.cpp:130-137; coarse buffers before the prolongation).
// "adaptive" makes the mesh adaptive (numlevel 3, derefine_count 2, derivative_order_1 on U
// component 3 with the VIBE tolerances 0.5 / 0.2): Refinement::Tag closes the last stage
// (burgers_driver.cpp:139-145) and every cycle ends with
// LoadBalancingAndAdaptiveMeshRefinement and, on a change, FillDerived and the time step of
// the new blocks (driver.cpp:110-131).
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <limits>
#include <string>
#include <vector>

#include "parthenon_amd.hpp"

using namespace parthenon;

namespace {

struct Integrator {
  int nstages = 2;
  Real beta[2] = {1.0, 0.5};
  const char *stage_name[3] = {"base", "1", "base"};
  Real dt = std::numeric_limits<Real>::max();
};

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s ncycles out.bin [sphere_radius] [open|periodic] [adaptive]\n",
                 argv[0]);
    return 2;
  }
  const int ncycles = std::atoi(argv[1]);
  const Real radius = argc > 3 ? std::atof(argv[3]) : 0.0;
  const bool open = argc > 4 && std::string(argv[4]) == "open";
  const bool adaptive = argc > 5 && std::string(argv[5]) == "adaptive";
  const Real cfl = 0.4;
  try {
    // burgers_package::Initialize: U (3 + 4 scalars) and the OneCopy derived field; with a
    // radius, two levels of static refinement on the blocks meeting a centred sphere
    Mesh mesh({64, 64, 64}, {16, 16, 16}, 4, {-0.5, -0.5, -0.5}, {0.5, 0.5, 0.5},
              {open ? 0 : 1, 1, open ? 0 : 1});
    if (open) mesh.SetBoundaryCondition(BoundaryFace::inner_x1, "reflecting");
    if (radius > 0.0) mesh.AddRefinementSphere({0.0, 0.0, 0.0}, radius, 2);
    mesh.AddField("U", 7, Metadata::FillGhost | Metadata::Independent | Metadata::WithFluxes);
    mesh.AddField("derived", 1, Metadata::OneCopy);
    if (adaptive) {
      mesh.SetAdaptive(3, 2);
      mesh.AddRefinementCriterion("derivative_order_1", "U", 3, 0.5, 0.2);
    }
    mesh.Initialize();
    auto &base = mesh.GetOrAdd("base", 0);
    if (pth_burgers_pgen_sinusoid(base->handle(), 0) != PTH_COMPLETE)
      throw std::runtime_error(pth_last_error());
    // Mesh::Initialize after the problem generator: CommunicateBoundaries, FillDerived
    SendBoundaryBuffers(base);
    ReceiveBoundaryBuffers(base);
    SetBoundaries(base);
    if (mesh.multilevel()) {
      ApplyBoundaryConditionsOnCoarseOrFineMD(base, true);
      ProlongateBoundaries(base);
    }
    ApplyBoundaryConditionsMD(base);
    burgers_package::CalculateDerived(base.get());
    Integrator integ;
    Real time = 0.0;
    integ.dt = std::min(integ.dt, burgers_package::EstimateTimestepMesh(base.get(), cfl));

    for (int cycle = 0; cycle < ncycles; ++cycle) {
      Real new_dt = std::numeric_limits<Real>::max();
      for (int stage = 1; stage <= integ.nstages; ++stage) {
        const Real beta = integ.beta[stage - 1];
        const Real dt = integ.dt;
        auto &mbase = mesh.GetOrAdd("base", 0);
        auto &mc0 = mesh.GetOrAdd(integ.stage_name[stage - 1], 0);
        auto &mc1 = mesh.GetOrAdd(integ.stage_name[stage], 0);
        auto &mdudt = mesh.GetOrAdd("dUdt", 0);
        constexpr auto any = BoundaryType::any;
        constexpr auto local = BoundaryType::local;
        constexpr auto nonlocal = BoundaryType::nonlocal;
        TaskList tl;
        tl.AddTask([&] { return StartReceiveBoundBufs<any>(mc1); });
        tl.AddTask([&] { return StartReceiveFluxCorrections(mc0); });
        tl.AddTask([&] { return burgers_package::CalculateFluxes(mc0.get()); });
        tl.AddTask([&] { return LoadAndSendFluxCorrections(mc0); });
        tl.AddTask([&] { return ReceiveFluxCorrections(mc0); });
        tl.AddTask([&] { return SetFluxCorrections(mc0); });
        tl.AddTask([&] { return Update::FluxDivergence(mc0.get(), mdudt.get()); });
        tl.AddTask([&] { return Update::AverageIndependentData(mc0.get(), mbase.get(), beta); });
        tl.AddTask([&] {
          return Update::UpdateIndependentData(mc0.get(), mdudt.get(), beta * dt, mc1.get());
        });
        tl.AddTask([&] { return SendBoundBufs<nonlocal>(mc1); });
        tl.AddTask([&] { return SendBoundBufs<local>(mc1); });
        tl.AddTask([&] { return ReceiveBoundBufs<local>(mc1); });
        tl.AddTask([&] { return SetBounds<local>(mc1); });
        tl.AddTask([&] { return ReceiveBoundBufs<nonlocal>(mc1); });
        tl.AddTask([&] { return SetBounds<nonlocal>(mc1); });
        // the library exchange order (AddBoundaryExchangeTasks): prolongation after the sets
        if (mesh.multilevel()) {
          tl.AddTask([&] { return ApplyBoundaryConditionsOnCoarseOrFineMD(mc1, true); });
          tl.AddTask([&] { return ProlongateBounds<any>(mc1); });
        }
        tl.AddTask([&] { return ApplyBoundaryConditionsMD(mc1); });
        tl.AddTask([&] {
          burgers_package::CalculateDerived(mc1.get());
          return TaskStatus::complete;
        });
        if (stage == integ.nstages)
          tl.AddTask([&] {
            new_dt = burgers_package::EstimateTimestepMesh(mc1.get(), cfl);
            return TaskStatus::complete;
          });
        if (stage == integ.nstages && adaptive)
          tl.AddTask([&] { return Refinement::Tag(mc1.get()); });
        tl.Execute();
      }
      time += integ.dt;
      if (adaptive && mesh.LoadBalancingAndAdaptiveMeshRefinement()) {
        // InitializeBlockTimeSteps on the new blocks (FillDerived ran in the remesh)
        auto &mb = mesh.GetOrAdd("base", 0);
        burgers_package::CalculateDerived(mb.get());
        new_dt = burgers_package::EstimateTimestepMesh(mb.get(), cfl);
      }
      // SetGlobalTimeStep: growth limited to 2x
      integ.dt = std::min(2.0 * integ.dt, new_dt);
    }
    std::vector<Real> U(static_cast<size_t>(mesh.nbtotal()) * 7 * 24 * 24 * 24);
    base->Download("U", U);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) throw std::runtime_error("cannot open output");
    std::fwrite(U.data(), sizeof(Real), U.size(), f);
    std::fwrite(&integ.dt, sizeof(Real), 1, f);
    std::fwrite(&time, sizeof(Real), 1, f);
    std::fclose(f);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "burgers_task_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
