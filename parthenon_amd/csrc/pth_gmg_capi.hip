// C ABI of the geometric-multigrid Poisson solver and of its serving loop.
//
// pth_gmg: one solver for one N^3 problem, the reference's MGSolver::AddTasks iteration
// (src/solvers/mg_solver.hpp:119-161: V-cycle, residual, stop test) as a host loop over the
// batched kernels of pth_gmg.hip.
//
// pth_gmg_server: continuous batching of independent solves. Requests are admitted when their
// arrival time has passed and the device memory budget and the slot limit allow; every outer
// iteration runs the same V-cycle stage of every active solve as one launch; a converged
// solve retires at once and its memory returns to a pool keyed by grid size, so admission
// between iterations needs no device allocation once the pool is warm. Each entry of a
// batched launch computes exactly what it computes alone, so every result is bit-identical
// alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "pth_gmg.hpp"

namespace pth {
namespace {

struct GmgHandle {
  GmgParams P;
  int N = 0;
  std::unique_ptr<GmgSolver1<float>> f;
  std::unique_ptr<GmgSolver1<double>> d;
  template <typename F>
  pth_status Do(F fn) {
    return f ? fn(*f) : fn(*d);
  }
};

// ---- serving ------------------------------------------------------------------------------
struct Request {
  int id, n;
  std::uint64_t seed;
  double arrival;
  // results
  double admit = -1, done = -1, rel = 0;
  int iters = 0;
  bool finished = false;
  std::vector<double> solution;
};

template <typename T>
struct Server {
  GmgParams P;
  double tol;
  int maxit, max_active;
  std::int64_t budget;           // cap on bytes held (<= 0: none beyond free HBM)
  std::int64_t reserve = 1LL << 30;  // free HBM left unallocated by admissions
  bool keep;
  hipStream_t s = nullptr;
  GmgBatch<T> batch;
  std::vector<Request> reqs;
  std::map<int, std::vector<GmgProblem<T> *>> pool;  // free problems by grid size
  std::int64_t held = 0;                             // bytes in active + pooled problems
  int iterations = 0, max_batch = 0;
  double busy_s = 0.0;
  ~Server() {
    for (auto &kv : pool)
      for (auto *p : kv.second) {
        p->Free();
        delete p;
      }
    batch.Free();
    if (s) (void)hipStreamDestroy(s);
  }
  pth_status Init() {
    PTH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return batch.Init(s);
  }
  // Admission by memory: the device's free HBM (hipMemGetInfo) must hold the problem and the
  // reserve, and the bytes held must stay within the cap when one is set.
  bool Fits(std::int64_t need) const {
    if (budget > 0 && held + need > budget) return false;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
    return static_cast<std::int64_t>(free_b) >= need + reserve;
  }
  GmgProblem<T> *Take(int n) {
    auto &v = pool[n];
    if (!v.empty()) {
      GmgProblem<T> *p = v.back();
      v.pop_back();
      return p;
    }
    const std::int64_t need = GmgProblem<T>::Bytes(n);
    if (!Fits(need)) {  // evict pooled problems of other sizes first
      for (auto &kv : pool)
        while (!kv.second.empty() && !Fits(need)) {
          GmgProblem<T> *q = kv.second.back();
          kv.second.pop_back();
          held -= GmgProblem<T>::Bytes(q->N);
          q->Free();
          delete q;
        }
      if (!Fits(need)) return nullptr;
    }
    auto *p = new GmgProblem<T>();
    if (p->Allocate(n, GmgBottom<T>::kN) != PTH_COMPLETE) {
      p->Free();
      delete p;
      return nullptr;
    }
    held += need;
    return p;
  }

  pth_status Run(bool realtime) {
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    std::vector<int> order(reqs.size());
    for (size_t q = 0; q < reqs.size(); ++q) order[q] = static_cast<int>(q);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
    size_t next = 0;
    struct Active {
      int r;
      GmgProblem<T> *p;
      double bnorm;
      bool corrected;
    };
    std::vector<Active> act;
    std::vector<int> waiting;  // arrived, not admitted
    while (next < order.size() || !waiting.empty() || !act.empty()) {
      const double t = realtime ? now() : 1e300;
      while (next < order.size() && reqs[order[next]].arrival <= t) waiting.push_back(order[next++]);
      // admission (arrival order) while slots and memory allow
      std::vector<Active> fresh;
      while (!waiting.empty() && static_cast<int>(act.size() + fresh.size()) < max_active) {
        Request &r = reqs[waiting.front()];
        GmgProblem<T> *p = Take(r.n);
        if (!p) break;  // memory: wait for a retirement
        waiting.erase(waiting.begin());
        if (batch.SetBlobs(*p, r.seed) != PTH_COMPLETE) return PTH_FAIL;
        if (batch.Reset(*p) != PTH_COMPLETE) return PTH_FAIL;
        r.admit = now();
        fresh.push_back({r.id, p, 0.0, false});
      }
      if (!fresh.empty()) {
        std::vector<GmgProblem<T> *> ps;
        for (auto &a : fresh) ps.push_back(a.p);
        std::vector<double> nb(ps.size());
        if (batch.RhsNorms(ps, nb.data()) != PTH_COMPLETE) return PTH_FAIL;
        for (size_t q = 0; q < fresh.size(); ++q) fresh[q].bnorm = nb[q];
        act.insert(act.end(), fresh.begin(), fresh.end());
      }
      if (act.empty()) {
        if (!waiting.empty()) {
          SetError("GMG server: a request does not fit in free device memory (or the cap)");
          return PTH_FAIL;
        }
        if (next < order.size() && realtime) {
          const double wait = reqs[order[next]].arrival - now();
          if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
        }
        continue;
      }
      // one outer iteration of every active solve
      const double tb = now();
      std::vector<GmgProblem<T> *> ps;
      std::vector<char> corr;
      for (auto &a : act) {
        ps.push_back(a.p);
        corr.push_back(a.corrected ? 1 : 0);
      }
      std::vector<double> rn(ps.size());
      if (batch.Residual(ps, corr, rn.data()) != PTH_COMPLETE) return PTH_FAIL;
      std::vector<Active> keep_act;
      std::vector<GmgProblem<T> *> cyc;
      for (size_t q = 0; q < act.size(); ++q) {
        Active &a = act[q];
        Request &r = reqs[a.r];
        const double rel = a.bnorm > 0 ? rn[q] / a.bnorm : rn[q];
        r.rel = rel;
        if (rel < tol || r.iters >= maxit) {
          if (keep) {
            const std::int64_t c = static_cast<std::int64_t>(r.n) * r.n * r.n;
            r.solution.resize(c);
            PTH_HIP(hipMemcpyAsync(r.solution.data(), a.p->u[a.p->cur], sizeof(double) * c,
                                   hipMemcpyDeviceToHost, s));
            PTH_HIP(hipStreamSynchronize(s));
          }
          r.done = now();
          r.finished = true;
          pool[r.n].push_back(a.p);
          continue;
        }
        ++r.iters;
        a.corrected = true;
        keep_act.push_back(a);
        cyc.push_back(a.p);
      }
      batch.jacobi_columns = P.columns;
      for (int c = 0; c < P.inner && !cyc.empty(); ++c)
        if (batch.Cycle(cyc, P.omega, P.pre, P.post, c > 0) != PTH_COMPLETE) return PTH_FAIL;
      max_batch = std::max(max_batch, static_cast<int>(act.size()));
      act.swap(keep_act);
      ++iterations;
      busy_s += now() - tb;
    }
    PTH_HIP(hipStreamSynchronize(s));
    return PTH_COMPLETE;
  }
};

struct ServerHandle {
  std::unique_ptr<Server<float>> f;
  std::unique_ptr<Server<double>> d;
  template <typename F>
  auto Do(F fn) {
    return f ? fn(*f) : fn(*d);
  }
};

}  // namespace
}  // namespace pth

using namespace pth;

struct pth_gmg : GmgHandle {};
struct pth_gmg_server : ServerHandle {};

extern "C" {

pth_gmg *pth_gmg_create(int n, int precision, double omega, int pre, int post) {
  if (n < 1 || (precision != 0 && precision != 1) || pre < 0 || post < 0) {
    SetError("pth_gmg_create: bad arguments");
    return nullptr;
  }
  auto *h = new pth_gmg();
  h->P = {precision, omega, pre, post};
  h->N = n;
  pth_status st;
  if (precision == 0) {
    h->f.reset(new GmgSolver1<float>());
    st = h->f->Init(n);
  } else {
    h->d.reset(new GmgSolver1<double>());
    st = h->d->Init(n);
  }
  if (st != PTH_COMPLETE) {
    delete h;
    return nullptr;
  }
  return h;
}

void pth_gmg_destroy(pth_gmg *h) { delete h; }

int pth_gmg_levels(const pth_gmg *h) { return static_cast<int>(GmgLevelSizes(h->N).size()); }

pth_status pth_gmg_set_rhs_blobs(pth_gmg *h, uint64_t seed) {
  return h->Do([&](auto &S) {
    if (S.batch.SetBlobs(S.p, seed) != PTH_COMPLETE) return PTH_FAIL;
    if (S.batch.Reset(S.p) != PTH_COMPLETE) return PTH_FAIL;
    return S.Norm();
  });
}

pth_status pth_gmg_set_rhs(pth_gmg *h, const double *host, size_t n) {
  const size_t c = static_cast<size_t>(h->N) * h->N * h->N;
  if (n != c) {
    SetError("pth_gmg_set_rhs: size mismatch");
    return PTH_FAIL;
  }
  return h->Do([&](auto &S) {
    PTH_HIP(hipMemcpyAsync(S.p.b64, host, sizeof(double) * c, hipMemcpyHostToDevice, S.s));
    if (S.batch.Reset(S.p) != PTH_COMPLETE) return PTH_FAIL;
    return S.Norm();
  });
}

pth_status pth_gmg_get_rhs(pth_gmg *h, double *host, size_t n) {
  const size_t c = static_cast<size_t>(h->N) * h->N * h->N;
  if (n != c) {
    SetError("pth_gmg_get_rhs: size mismatch");
    return PTH_FAIL;
  }
  return h->Do([&](auto &S) {
    PTH_HIP(hipMemcpyAsync(host, S.p.b64, sizeof(double) * c, hipMemcpyDeviceToHost, S.s));
    PTH_HIP(hipStreamSynchronize(S.s));
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_get_solution(pth_gmg *h, double *host, size_t n) {
  const size_t c = static_cast<size_t>(h->N) * h->N * h->N;
  if (n != c) {
    SetError("pth_gmg_get_solution: size mismatch");
    return PTH_FAIL;
  }
  return h->Do([&](auto &S) {
    PTH_HIP(hipMemcpyAsync(host, S.p.u[S.p.cur], sizeof(double) * c, hipMemcpyDeviceToHost,
                           S.s));
    PTH_HIP(hipStreamSynchronize(S.s));
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_reset(pth_gmg *h) {
  return h->Do([&](auto &S) {
    if (S.batch.Reset(S.p) != PTH_COMPLETE) return PTH_FAIL;
    PTH_HIP(hipStreamSynchronize(S.s));
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_set_inner_cycles(pth_gmg *h, int k) {
  if (k < 1) {
    SetError("pth_gmg_set_inner_cycles: k must be >= 1");
    return PTH_FAIL;
  }
  h->P.inner = k;
  return PTH_COMPLETE;
}

pth_status pth_gmg_set_column_sweeps(pth_gmg *h, int on) {
  h->P.columns = on != 0;
  return PTH_COMPLETE;
}

pth_status pth_gmg_solve(pth_gmg *h, double tol, int max_iters, int *iters, double *final_rel) {
  return h->Do([&](auto &S) { return S.Solve(h->P, tol, max_iters, iters, final_rel); });
}

int pth_gmg_history(const pth_gmg *h, double *out, int max) {
  const std::vector<double> &v = h->f ? h->f->history : h->d->history;
  const int n = std::min(static_cast<int>(v.size()), max);
  for (int q = 0; q < n; ++q) out[q] = v[q];
  return static_cast<int>(v.size());
}

pth_gmg_server *pth_gmg_server_create(int precision, double omega, int pre, int post, double tol,
                                      int max_iters, int64_t budget_bytes, int max_active,
                                      int keep_solutions) {
  if ((precision != 0 && precision != 1) || max_active < 1) {
    SetError("pth_gmg_server_create: bad arguments");
    return nullptr;
  }
  auto *h = new pth_gmg_server();
  auto setup = [&](auto &S) {
    S.P = {precision, omega, pre, post};
    S.tol = tol;
    S.maxit = max_iters;
    S.budget = budget_bytes;
    S.max_active = std::min(max_active, GmgBatch<float>::kMaxProblems);
    S.keep = keep_solutions != 0;
    return S.Init();
  };
  pth_status st;
  if (precision == 0) {
    h->f.reset(new Server<float>());
    st = setup(*h->f);
  } else {
    h->d.reset(new Server<double>());
    st = setup(*h->d);
  }
  if (st != PTH_COMPLETE) {
    delete h;
    return nullptr;
  }
  return h;
}

void pth_gmg_server_destroy(pth_gmg_server *h) { delete h; }

pth_status pth_gmg_server_submit(pth_gmg_server *h, int n, uint64_t seed, double arrival_s,
                                 int *id) {
  if (n < 1) {
    SetError("pth_gmg_server_submit: bad grid size");
    return PTH_FAIL;
  }
  return h->Do([&](auto &S) {
    Request r;
    r.id = static_cast<int>(S.reqs.size());
    r.n = n;
    r.seed = seed;
    r.arrival = arrival_s;
    S.reqs.push_back(std::move(r));
    if (id) *id = S.reqs.back().id;
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_server_set_inner_cycles(pth_gmg_server *h, int k) {
  if (k < 1) {
    SetError("pth_gmg_server_set_inner_cycles: k must be >= 1");
    return PTH_FAIL;
  }
  return h->Do([&](auto &S) {
    S.P.inner = k;
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_server_set_hbm_reserve(pth_gmg_server *h, int64_t bytes) {
  if (bytes < 0) {
    SetError("pth_gmg_server_set_hbm_reserve: bytes must be >= 0");
    return PTH_FAIL;
  }
  return h->Do([&](auto &S) {
    S.reserve = bytes;
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_server_set_column_sweeps(pth_gmg_server *h, int on) {
  return h->Do([&](auto &S) {
    S.P.columns = on != 0;
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_server_run(pth_gmg_server *h, int realtime) {
  return h->Do([&](auto &S) { return S.Run(realtime != 0); });
}

pth_status pth_gmg_server_result(const pth_gmg_server *h, int id, double out[5], int *iters) {
  auto *hh = const_cast<pth_gmg_server *>(h);
  return hh->Do([&](auto &S) {
    if (id < 0 || id >= static_cast<int>(S.reqs.size()) || !S.reqs[id].finished) {
      SetError("pth_gmg_server_result: no finished request with this id");
      return PTH_FAIL;
    }
    const Request &r = S.reqs[id];
    out[0] = r.arrival;
    out[1] = r.admit;
    out[2] = r.done;
    out[3] = r.done - r.arrival;
    out[4] = r.rel;
    if (iters) *iters = r.iters;
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_server_solution(const pth_gmg_server *h, int id, double *host, size_t n) {
  auto *hh = const_cast<pth_gmg_server *>(h);
  return hh->Do([&](auto &S) {
    if (id < 0 || id >= static_cast<int>(S.reqs.size()) || S.reqs[id].solution.size() != n) {
      SetError("pth_gmg_server_solution: no kept solution of this size");
      return PTH_FAIL;
    }
    std::copy(S.reqs[id].solution.begin(), S.reqs[id].solution.end(), host);
    return PTH_COMPLETE;
  });
}

pth_status pth_gmg_server_stats(const pth_gmg_server *h, int64_t out[3], double *busy_s) {
  auto *hh = const_cast<pth_gmg_server *>(h);
  return hh->Do([&](auto &S) {
    out[0] = S.iterations;
    out[1] = S.max_batch;
    out[2] = S.held;
    if (busy_s) *busy_s = S.busy_s;
    return PTH_COMPLETE;
  });
}

}  // extern "C"
