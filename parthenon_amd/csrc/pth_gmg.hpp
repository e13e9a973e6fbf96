// Host-side structures of the geometric-multigrid Poisson solver (pth_gmg.hip).
#pragma once
#include <cstdint>
#include <vector>

#include "pth_internal.hpp"

namespace pth {

// One (problem, level) entry of a batched launch. Level arrays are n^3 cells, [k][j][i].
template <typename T>
struct GmgEntry {
  T *x, *t, *b;        // this level: solution (correction), Jacobi scratch, right-hand side
  T *xc, *bc;          // next coarser level's solution and right-hand side (or null)
  double *u_old, *u_new;  // finest level FP64 solution, ping-pong (outer step only)
  const double *b64;      // finest level FP64 right-hand side
  double *partial;        // per-workgroup partial sums of the outer step
  int n, nc;              // cells per side on this and the next coarser level
  int wg_begin;           // first workgroup of this entry in the batched grid
  int warm;               // bottom solve: start from x instead of 0
};

// The entries of one launch, passed as a kernel argument: inline for up to kGmgInline
// entries (no copy before the launch), else in device memory (dev, staged by GmgBatch).
constexpr int kGmgInline = 8;
template <typename T>
struct GmgArgs {
  GmgEntry<T> e[kGmgInline];
  const GmgEntry<T> *dev;
  int ne;
};

// largest level handled inside one workgroup in LDS (three n^3 arrays of T plus the coarser
// levels below the 64 KiB static LDS limit)
template <typename T>
struct GmgBottom;
template <>
struct GmgBottom<float> {
  static constexpr int kN = 16;
};
template <>
struct GmgBottom<double> {
  static constexpr int kN = 8;
};

template <typename T>
struct GmgProblem {
  struct Arrays {
    T *x, *t, *b;
  };
  int N = 0;
  std::vector<int> sizes;  // cells per side of every level, finest first
  int nbottom = 0;         // first level solved in LDS
  std::vector<Arrays> lv;
  T *levels = nullptr;     // one allocation for all levels
  double *u[2] = {nullptr, nullptr};
  double *b64 = nullptr, *partial = nullptr;
  int cur = 0;             // u[cur] is the current solution
  static std::int64_t Bytes(int N);  // device bytes one problem holds
  pth_status Allocate(int N, int bottom);
  void Free();
  GmgEntry<T> Entry(int l) const;
};

template <typename T>
struct GmgBatch {
  static constexpr int kStageCap = 8192;
  static constexpr int kMaxProblems = 1024;
  hipStream_t stream = nullptr;
  GmgEntry<T> *stage_host = nullptr, *stage_dev = nullptr;
  int stage_used = 0;
  double *norms_host = nullptr, *norms_dev = nullptr;
  pth_status Init(hipStream_t s);
  void Free();
  // per_wg: cells one workgroup of the launch covers
  pth_status Stage(const std::vector<GmgEntry<T>> &es, bool coarse_cells, GmgArgs<T> *args,
                   int *nwg, int per_wg = 256);
  int jacobi_columns = 1;  // Jacobi sweeps by z columns where the level size allows
  pth_status RhsNorms(const std::vector<GmgProblem<T> *> &ps, double *out);
  pth_status Residual(const std::vector<GmgProblem<T> *> &ps, const std::vector<char> &correct,
                      double *out);
  // warm: the finest level starts from its current x (a further inner cycle on the same
  // right-hand side) instead of 0
  pth_status Cycle(const std::vector<GmgProblem<T> *> &ps, double omega, int pre, int post,
                   bool warm = false);
  pth_status SetBlobs(GmgProblem<T> &p, std::uint64_t seed);
  pth_status Reset(GmgProblem<T> &p);
};

struct GmgParams {
  int precision;  // 0: FP32 inner cycle (mixed), 1: FP64
  double omega;
  int pre, post;
  int inner = 1;    // V-cycles per FP64 refinement step
  int columns = 1;  // Jacobi sweeps by z columns where the level size allows
};

// One solve of one problem: the MGSolver::AddTasks iteration (mg_solver.hpp:119-161) as a host
// loop over the batched kernels: residual, stop test, V-cycle(s).
template <typename T>
struct GmgSolver1 {
  GmgProblem<T> p;
  GmgBatch<T> batch;
  hipStream_t s = nullptr;
  double bnorm = 0.0;
  std::vector<double> history;
  ~GmgSolver1() {
    p.Free();
    batch.Free();
    if (s) (void)hipStreamDestroy(s);
  }
  pth_status Init(int N) {
    PTH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    if (batch.Init(s) != PTH_COMPLETE) return PTH_FAIL;
    return p.Allocate(N, GmgBottom<T>::kN);
  }
  pth_status Norm() { return batch.RhsNorms({&p}, &bnorm); }
  // relative residual ||b - A u|| / ||b|| (||r|| when b = 0) below tol ends the loop once
  // min_cycles V-cycles have run; max_iters cycles at most
  pth_status Solve(const GmgParams &P, double tol, int maxit, int *iters, double *rel,
                   int min_cycles = 0) {
    history.clear();
    std::vector<GmgProblem<T> *> ps = {&p};
    std::vector<char> corr = {0};
    int it = 0;
    for (;; ++it) {
      double rn;
      if (batch.Residual(ps, corr, &rn) != PTH_COMPLETE) return PTH_FAIL;
      const double r = bnorm > 0 ? rn / bnorm : rn;
      history.push_back(r);
      if ((r < tol && it >= min_cycles) || it >= maxit) break;
      batch.jacobi_columns = P.columns;
      for (int c = 0; c < P.inner; ++c)
        if (batch.Cycle(ps, P.omega, P.pre, P.post, c > 0) != PTH_COMPLETE) return PTH_FAIL;
      corr[0] = 1;
    }
    if (iters) *iters = it;
    if (rel) *rel = history.back();
    return PTH_COMPLETE;
  }
};

void GmgBlobs(std::uint64_t seed, double out[80]);
std::vector<int> GmgLevelSizes(int N);

}  // namespace pth
