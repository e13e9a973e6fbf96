// The poisson_gmg example driven by one of the reference's own input decks (parthinput format,
// e.g. tst/regression/test_suites/poisson_gmg/parthinput.poisson), run through
// include/parthenon_amd.hpp with the task collection of PoissonDriver::MakeTaskCollection
// (example/poisson_gmg/poisson_driver.cpp:64-133) and the MG solver.
//
// Read from the deck: <parthenon/mesh> nx1..3, x1min..x3max, ix1_bc..ox3_bc (periodic stays
// periodic; every other face gets the package's FixedFace conditions, which it registers by
// hand, poisson_package.cpp:39-78), <parthenon/meshblock> nx1..3, <parthenon/static_refinementN>
// boxes with their logical level, <poisson> use_exact_rhs, x0 y0 z0 radius, interior_D /
// exterior_D (the face field D), diagonal_alpha, <poisson/solver_params>
// max_iterations, residual_tolerance, smoother. do_FAS is read and not needed (the correction
// scheme gives the same iterates for this linear operator); the BiCGSTAB keys are ignored.
// The problem generator is parthenon_app_inputs.cpp:32-82: rhs = 1 inside the radius,
// exact = -exp(-10 r^2) (r over the active dimensions; x3 from the x3 centre), D on each face
// interior_D inside the two boxes of inside_region and exterior_D elsewhere (x3 of the face
// from the x3 coordinates, where the reference reads X<2>).
//
// usage: poisson_deck_driver <deck> <out.bin>
// writes doubles: final RMS residual, final RMS error (or -1), V-cycles, blocks, cells per
// block, then the interior cells of u, rhs and exact, block order then [k][j][i].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "parthenon_amd.hpp"

using namespace parthenon;

namespace poisson_package {
PARTHENON_AMD_VARIABLE(poisson, u);
PARTHENON_AMD_VARIABLE(poisson, rhs);
PARTHENON_AMD_VARIABLE(poisson, exact);
PARTHENON_AMD_VARIABLE(poisson, D);
}  // namespace poisson_package

namespace {

// <block> key = value, '#' comments (src/parameter_input.cpp's format, the parts decks use)
using Deck = std::map<std::string, std::map<std::string, std::string>>;

std::string Trim(const std::string &s) {
  const size_t a = s.find_first_not_of(" \t\r"), b = s.find_last_not_of(" \t\r");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

Deck Read(const std::string &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot read deck " + path);
  Deck d;
  std::string line, block;
  while (std::getline(in, line)) {
    line = Trim(line.substr(0, line.find('#')));
    if (line.empty()) continue;
    if (line.front() == '<') {
      block = Trim(line.substr(1, line.find('>') - 1));
      continue;
    }
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    d[block][Trim(line.substr(0, eq))] = Trim(line.substr(eq + 1));
  }
  return d;
}

std::string Get(const Deck &d, const std::string &b, const std::string &k,
                const std::string &def) {
  auto ib = d.find(b);
  if (ib == d.end()) return def;
  auto ik = ib->second.find(k);
  return ik == ib->second.end() ? def : ik->second;
}
Real GetReal(const Deck &d, const std::string &b, const std::string &k, Real def) {
  return std::atof(Get(d, b, k, std::to_string(def)).c_str());
}
int GetInt(const Deck &d, const std::string &b, const std::string &k, int def) {
  return std::atoi(Get(d, b, k, std::to_string(def)).c_str());
}
bool GetBool(const Deck &d, const std::string &b, const std::string &k, bool def) {
  const std::string v = Get(d, b, k, def ? "true" : "false");
  return v == "true" || v == "1";
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s deck out.bin\n", argv[0]);
    return 2;
  }
  using namespace poisson_package;
  try {
    const Deck deck = Read(argv[1]);
    const std::string M = "parthenon/mesh", B = "parthenon/meshblock", P = "poisson",
                      S = "poisson/solver_params";
    std::array<int, 3> nx, nb;
    std::array<Real, 3> lo, hi;
    std::array<int, 3> per;
    for (int d = 0; d < 3; ++d) {
      const std::string n = std::to_string(d + 1);
      nx[d] = GetInt(deck, M, "nx" + n, 1);
      nb[d] = GetInt(deck, B, "nx" + n, nx[d]);
      lo[d] = GetReal(deck, M, "x" + n + "min", 0.0);
      hi[d] = GetReal(deck, M, "x" + n + "max", 1.0);
      per[d] = Get(deck, M, "ix" + n + "_bc", "outflow") == "periodic" ? 1 : 0;
    }
    const int ndim = nx[2] > 1 ? 3 : nx[1] > 1 ? 2 : 1;
    const int ng = GetInt(deck, M, "nghost", 2);
    const Real interior_D = GetReal(deck, P, "interior_D", 1.0),
               exterior_D = GetReal(deck, P, "exterior_D", 1.0);
    const bool use_exact_rhs = GetBool(deck, P, "use_exact_rhs", false);
    const Real x0 = GetReal(deck, P, "x0", 0.0), y0 = GetReal(deck, P, "y0", 0.0),
               z0 = GetReal(deck, P, "z0", 0.0), radius0 = GetReal(deck, P, "radius", 0.1);

    Mesh mesh(nx, nb, ng, lo, hi, per);
    for (int f = 0; f < 2 * ndim; ++f)
      if (!per[f / 2]) mesh.SetBoundaryCondition(static_cast<BoundaryFace>(f), "fixed_face");
    const int root = pth_mesh_root_level(mesh.handle());
    for (int r = 0;; ++r) {
      const std::string blk = "parthenon/static_refinement" + std::to_string(r);
      if (!deck.count(blk)) break;
      std::array<Real, 6> box;
      for (int d = 0; d < 3; ++d) {
        const std::string n = std::to_string(d + 1);
        box[2 * d] = GetReal(deck, blk, "x" + n + "min", lo[d]);
        box[2 * d + 1] = GetReal(deck, blk, "x" + n + "max", hi[d]);
      }
      const int level = GetInt(deck, blk, "level", root) - root;  // logical -> above base
      if (level > 0) mesh.AddStaticRefinement(box, level);
    }
    mesh.AddField(u::name(), 1, Metadata::FillGhost | Metadata::Independent);
    // the package registers ProlongateSharedLinear for u (poisson_package.cpp:122-126)
    mesh.RegisterRefinementOps(u::name(), Mesh::ProlongationOp::SharedLinear);
    mesh.AddField(rhs::name(), 1, Metadata::OneCopy);
    mesh.AddField(exact::name(), 1, Metadata::OneCopy);
    mesh.AddField(D::name(), 3, Metadata::OneCopy);  // component d: D on the lower face along d
    mesh.Initialize();

    // ProblemGenerator (parthenon_app_inputs.cpp:32-82) on every cell, ghosts included
    int shape[12];
    if (pth_mesh_index_shapes(mesh.handle(), shape) != PTH_COMPLETE)
      throw std::runtime_error(pth_last_error());
    const int n[3] = {shape[0], shape[1], shape[2]}, g[3] = {shape[3], shape[4], shape[5]};
    const int e[3] = {n[0] + 2 * g[0], n[1] + 2 * g[1], n[2] + 2 * g[2]};
    const size_t cells = static_cast<size_t>(e[0]) * e[1] * e[2];
    const int nblocks = mesh.nbtotal();
    std::vector<Real> hr(nblocks * cells), hu(hr.size(), 0.0), he(hr.size()),
        hd(3 * hr.size());
    auto inside_region = [ndim](Real x, Real y, Real z) {
      bool inside1 = (x < -0.25) && (x > -0.75);
      if (ndim > 1) inside1 = inside1 && (y < 0.5) && (y > -0.5);
      if (ndim > 2) inside1 = inside1 && (z < 0.25) && (z > -0.25);
      bool inside2 = (x < 0.5) && (x > -0.75);
      if (ndim > 1) inside2 = inside2 && (y < -0.25) && (y > -0.75);
      if (ndim > 2) inside2 = inside2 && (z < 0.25) && (z > -0.25);
      return inside1 || inside2;
    };
    for (int b = 0; b < nblocks; ++b) {
      int64_t loc[4];
      double x[6];
      if (pth_mesh_block_info(mesh.handle(), b, loc, x) != PTH_COMPLETE)
        throw std::runtime_error(pth_last_error());
      Real dx[3];
      for (int d = 0; d < 3; ++d) dx[d] = (x[3 + d] - x[d]) / n[d];
      for (int k = 0; k < e[2]; ++k)
        for (int j = 0; j < e[1]; ++j)
          for (int i = 0; i < e[0]; ++i) {
            const Real x1 = x[0] + (i - g[0] + 0.5) * dx[0], x2 = x[1] + (j - g[1] + 0.5) * dx[1],
                       x3 = x[2] + (k - g[2] + 0.5) * dx[2];
            Real rad = (x1 - x0) * (x1 - x0);
            if (ndim > 1) rad += (x2 - y0) * (x2 - y0);
            if (ndim > 2) rad += (x3 - z0) * (x3 - z0);
            rad = std::sqrt(rad);
            const size_t c = b * cells + (static_cast<size_t>(k) * e[1] + j) * e[0] + i;
            hr[c] = rad < radius0 ? 1.0 : 0.0;
            he[c] = -std::exp(-10.0 * rad * rad);
            const Real x1f = x1 - 0.5 * dx[0], x2f = x2 - 0.5 * dx[1], x3f = x3 - 0.5 * dx[2];
            const size_t dc = b * 3 * cells + (static_cast<size_t>(k) * e[1] + j) * e[0] + i;
            hd[dc] = inside_region(x1f, x2, x3) ? interior_D : exterior_D;
            hd[dc + cells] = inside_region(x1, x2f, x3) ? interior_D : exterior_D;
            hd[dc + 2 * cells] = inside_region(x1, x2, x3f) ? interior_D : exterior_D;
          }
    }
    auto &base = mesh.GetOrAdd("base", 0);
    base->Upload(rhs::name(), hr);
    base->Upload(u::name(), hu);
    base->Upload(exact::name(), he);
    base->Upload(D::name(), hd);

    // the package's solver (poisson_package.cpp:80-98) with the deck's solver_params
    solvers::PoissonEquation eq;
    eq.D_field = D::name();
    eq.diagonal_alpha = GetReal(deck, P, "diagonal_alpha", 0.0);
    solvers::MGParams params;
    params.max_iters = GetInt(deck, S, "max_iterations", 1000);
    params.residual_tolerance = GetReal(deck, S, "residual_tolerance", 1.e-12);
    params.smoother = Get(deck, S, "smoother", "SRJ2");
    params.mixed_precision = false;
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
      auto get_rhs = none;
      if (use_exact_rhs) {
        auto copy_exact = tl.AddTask(get_rhs, TF(solvers::utils::CopyData<exact, u>), md);
        auto comm = AddBoundaryExchangeTasks<BoundaryType::any>(copy_exact, tl, md, true);
        get_rhs = eq.Ax<u, rhs>(tl, comm, md);
      }
      auto zero_u = tl.AddTask(get_rhs, TF(solvers::utils::SetToZero<u>), md);
      auto setup = mg_solver.AddSetupTasks(tl, zero_u, i, &mesh);
      auto solve = mg_solver.AddTasks(tl, setup, &mesh, i);
      auto keep = tl.AddTask(solve, [&]() {
        solution.resize(nblocks * cells);
        md->Download(u::name(), solution);
        return TaskStatus::complete;
      });
      if (use_exact_rhs) {
        auto diff = tl.AddTask(keep, TF(solvers::utils::AddFieldsAndStore<exact, u, u>), md,
                               1.0, -1.0);
        auto get_err = solvers::utils::DotProduct<u, u>(diff, tl, &err, md);
        tl.AddTask(get_err, [&, i]() {
          if (i != 0) return TaskStatus::complete;
          final_rms_error =
              std::sqrt(err.val / (static_cast<Real>(nblocks) * n[0] * n[1] * n[2]));
          std::printf("Final rms error: %e\n", final_rms_error);
          return TaskStatus::complete;
        });
      }
    }
    tc.Execute();
    const Real final_rms_residual = mg_solver.GetFinalResidual();
    const int iters = mg_solver.GetFinalIterations();
    std::printf("V-cycles: %d, final rms residual: %e\n", iters, final_rms_residual);

    std::vector<Real> hr2(nblocks * cells), he2(nblocks * cells);
    base->Download(rhs::name(), hr2);
    base->Download(exact::name(), he2);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f) throw std::runtime_error("cannot open output");
    const Real head[5] = {final_rms_residual, final_rms_error, static_cast<Real>(iters),
                          static_cast<Real>(nblocks), static_cast<Real>(n[0] * n[1] * n[2])};
    std::fwrite(head, sizeof(Real), 5, f);
    for (const std::vector<Real> *v : {&solution, &hr2, &he2})
      for (int b = 0; b < nblocks; ++b)
        for (int k = g[2]; k < g[2] + n[2]; ++k)
          for (int j = g[1]; j < g[1] + n[1]; ++j)
            std::fwrite(&(*v)[b * cells + (static_cast<size_t>(k) * e[1] + j) * e[0] + g[0]],
                        sizeof(Real), n[0], f);
    std::fclose(f);
  } catch (const std::exception &ex) {
    std::fprintf(stderr, "poisson_deck_driver: %s\n", ex.what());
    return 1;
  }
  return 0;
}
