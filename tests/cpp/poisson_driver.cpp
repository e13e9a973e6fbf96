// A C++ driver written against the reference's task-level names for the multigrid Poisson
// example, run through include/parthenon_amd.hpp: the task collection of
// PoissonDriver::MakeTaskCollection (example/poisson_gmg/poisson_driver.cpp:64-133) with the
// MG solver (MGSolver<u, rhs, PoissonEquation>::AddSetupTasks / AddTasks), the package's
// FixedFace conditions on every face (poisson_package.cpp:60-78) and the problem generator's
// exact solution u = -exp(-10 r^2) (parthenon_app_inputs.cpp:32-82; the radius in 3-D uses
// x3, where the reference reads X<2> for x3).
//
// usage: poisson_driver <N> <block> <out.bin> [mixed|fp64] [exact|sphere] [side] [D] [tol]
//   N^3 cells on the cube [-side/2, side/2]^3 (side 2 by default, the reference deck's
//   [-1, 1]) in block^3 meshblocks, nghost 2; "exact" (the default) solves A u = A exact
//   (use_exact_rhs, poisson_driver.cpp:86-97) and reports the RMS error against exact;
//   "sphere" solves with rhs = 1 inside r < 0.5 (the generator's default right-hand side).
// writes doubles: final RMS residual, final RMS error (or -1), V-cycles, then the global
// N^3 arrays ([k][j][i]) of the solution u, rhs and exact.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "parthenon_amd.hpp"

using namespace parthenon;

namespace poisson_package {
PARTHENON_AMD_VARIABLE(poisson, u);
PARTHENON_AMD_VARIABLE(poisson, rhs);
PARTHENON_AMD_VARIABLE(poisson, exact);
}  // namespace poisson_package

namespace {

struct Grid {
  int N, nb, ng, ent;
};

// interior cells of every block into a global [k][j][i] array (block positions from the mesh)
std::vector<Real> Assemble(const Mesh &mesh, const std::vector<Real> &bricks, const Grid &g) {
  std::vector<Real> out(static_cast<size_t>(g.N) * g.N * g.N);
  const size_t cells = static_cast<size_t>(g.ent) * g.ent * g.ent;
  for (int b = 0; b < mesh.nbtotal(); ++b) {
    int64_t loc[4];
    double x[6];
    if (pth_mesh_block_info(mesh.handle(), b, loc, x) != PTH_COMPLETE)
      throw std::runtime_error(pth_last_error());
    for (int k = 0; k < g.nb; ++k)
      for (int j = 0; j < g.nb; ++j)
        for (int i = 0; i < g.nb; ++i) {
          const size_t K = loc[3] * g.nb + k, J = loc[2] * g.nb + j, I = loc[1] * g.nb + i;
          out[(K * g.N + J) * g.N + I] =
              bricks[b * cells + ((static_cast<size_t>(k + g.ng) * g.ent + j + g.ng) * g.ent +
                                  i + g.ng)];
        }
  }
  return out;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 4) {
    std::fprintf(stderr,
                 "usage: %s N block out.bin [mixed|fp64] [exact|sphere] [side] [D] [tol]\n",
                 argv[0]);
    return 2;
  }
  const int N = std::atoi(argv[1]), nb = std::atoi(argv[2]);
  const bool fp64 = argc > 4 && std::string(argv[4]) == "fp64";
  const bool use_exact_rhs = !(argc > 5 && std::string(argv[5]) == "sphere");
  const Real side = argc > 6 ? std::atof(argv[6]) : 2.0;
  const Real Dcoef = argc > 7 ? std::atof(argv[7]) : 1.0;
  const Real tol = argc > 8 ? std::atof(argv[8]) : 1.e-12;
  const int ng = 2;  // the multigrid default nghost (src/parthenon_manager.cpp:141)
  using namespace poisson_package;
  try {
    Mesh mesh({N, N, N}, {nb, nb, nb}, ng, {-0.5 * side, -0.5 * side, -0.5 * side},
              {0.5 * side, 0.5 * side, 0.5 * side}, {0, 0, 0});
    for (int f = 0; f < 6; ++f)
      mesh.SetBoundaryCondition(static_cast<BoundaryFace>(f), "fixed_face");
    mesh.AddField(u::name(), 1, Metadata::FillGhost | Metadata::Independent);
    mesh.AddField(rhs::name(), 1, Metadata::OneCopy);
    mesh.AddField(exact::name(), 1, Metadata::OneCopy);
    mesh.Initialize();

    // ProblemGenerator: rhs, u = 0 and exact on every cell (ghosts included)
    const Grid g{N, nb, ng, nb + 2 * ng};
    const size_t cells = static_cast<size_t>(g.ent) * g.ent * g.ent;
    std::vector<Real> hr(mesh.nbtotal() * cells), hu(hr.size(), 0.0), he(hr.size());
    for (int b = 0; b < mesh.nbtotal(); ++b) {
      int64_t loc[4];
      double x[6];
      if (pth_mesh_block_info(mesh.handle(), b, loc, x) != PTH_COMPLETE)
        throw std::runtime_error(pth_last_error());
      const Real dx = (x[3] - x[0]) / nb;
      for (int k = 0; k < g.ent; ++k)
        for (int j = 0; j < g.ent; ++j)
          for (int i = 0; i < g.ent; ++i) {
            const Real x1 = x[0] + (i - ng + 0.5) * dx, x2 = x[1] + (j - ng + 0.5) * dx,
                       x3 = x[2] + (k - ng + 0.5) * dx;
            const Real rad = std::sqrt(x1 * x1 + x2 * x2 + x3 * x3);
            const size_t c = b * cells + (static_cast<size_t>(k) * g.ent + j) * g.ent + i;
            hr[c] = rad < 0.5 ? 1.0 : 0.0;
            he[c] = -std::exp(-10.0 * rad * rad);
          }
    }
    auto &base = mesh.GetOrAdd("base", 0);
    base->Upload(rhs::name(), hr);
    base->Upload(u::name(), hu);
    base->Upload(exact::name(), he);

    // the package's solver object (poisson_package.cpp:93-98)
    solvers::PoissonEquation eq;
    eq.D = Dcoef;
    solvers::MGParams params;
    params.max_iters = 40;
    params.residual_tolerance = tol;
    params.mixed_precision = !fp64;
    solvers::MGSolver<u, rhs, solvers::PoissonEquation> mg_solver(params, eq);

    // PoissonDriver::MakeTaskCollection (poisson_driver.cpp:64-133)
    AllReduce<Real> err;
    Real final_rms_error = -1.0;
    std::vector<Real> solution;
    TaskCollection tc;
    TaskID none;
    const int num_partitions = mesh.DefaultNumPartitions();
    TaskRegion &region = tc.AddRegion(num_partitions);
    for (int i = 0; i < num_partitions; ++i) {
      TaskList &tl = region[i];
      auto &md = mesh.GetOrAdd("base", i);

      // rhs <- A.exact, so that the solution of A.u = rhs is known
      auto get_rhs = none;
      if (use_exact_rhs) {
        auto copy_exact = tl.AddTask(get_rhs, TF(solvers::utils::CopyData<exact, u>), md);
        auto comm =
            AddBoundaryExchangeTasks<BoundaryType::any>(copy_exact, tl, md, mesh.multilevel());
        get_rhs = eq.Ax<u, rhs>(tl, comm, md);
      }

      // initial guess zero
      auto zero_u = tl.AddTask(get_rhs, TF(solvers::utils::SetToZero<u>), md);

      auto setup = mg_solver.AddSetupTasks(tl, zero_u, i, &mesh);
      auto solve = mg_solver.AddTasks(tl, setup, &mesh, i);

      // keep the solution before the error overwrites u
      auto keep = tl.AddTask(solve, [&]() {
        std::vector<Real> h(mesh.nbtotal() * cells);
        md->Download(u::name(), h);
        solution = Assemble(mesh, h, g);
        return TaskStatus::complete;
      });

      if (use_exact_rhs) {
        auto diff = tl.AddTask(keep, TF(solvers::utils::AddFieldsAndStore<exact, u, u>), md,
                               1.0, -1.0);
        auto get_err = solvers::utils::DotProduct<u, u>(diff, tl, &err, md);
        tl.AddTask(get_err, [&, i]() {
          if (i != 0) return TaskStatus::complete;
          final_rms_error = std::sqrt(err.val / (static_cast<Real>(N) * N * N));
          std::printf("Final rms error: %e\n", final_rms_error);
          return TaskStatus::complete;
        });
      }
    }
    tc.Execute();

    // PoissonDriver::Execute: the final residual for checking (poisson_driver.cpp:38-61)
    const Real final_rms_residual = mg_solver.GetFinalResidual();
    const int iters = mg_solver.GetFinalIterations();
    std::printf("V-cycles: %d, final rms residual: %e\n", iters, final_rms_residual);

    std::vector<Real> hr2(mesh.nbtotal() * cells), he2(hr2.size());
    base->Download(rhs::name(), hr2);
    base->Download(exact::name(), he2);
    const std::vector<Real> grhs = Assemble(mesh, hr2, g), gex = Assemble(mesh, he2, g);
    FILE *f = std::fopen(argv[3], "wb");
    if (!f) throw std::runtime_error("cannot open output");
    const Real head[3] = {final_rms_residual, final_rms_error, static_cast<Real>(iters)};
    std::fwrite(head, sizeof(Real), 3, f);
    std::fwrite(solution.data(), sizeof(Real), solution.size(), f);
    std::fwrite(grhs.data(), sizeof(Real), grhs.size(), f);
    std::fwrite(gex.data(), sizeof(Real), gex.size(), f);
    std::fclose(f);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "poisson_driver: %s\n", e.what());
    return 1;
  }
  return 0;
}
