// The multigrid Poisson solver over MeshData, its equation's operator and the solver field
// utilities, all on the bricks of a MeshData.
//
//   LinearKernel     CopyData / AddFieldsAndStore (src/solvers/solver_utils.hpp:151-222):
//                    out = wa * a + wb * b (or out = a) on every cell, ghosts included
//   ZeroKernel       SetToZero (solver_utils.hpp:224-250): interior cells
//   DotKernel        DotProductLocal (solver_utils.hpp:252-277): sum over interior cells and
//                    components of a * b; per-workgroup partials, then one workgroup adds them
//                    in a fixed order, so every call gives the same bits
//   PoissonAxKernel  PoissonEquation::Ax (example/poisson_gmg/poisson_equation.hpp:41-53) =
//                    CalculateFluxes (:131-184) + FluxMultiplyMatrix (:186-205) with a constant
//                    D: out = -alpha x + sum_d (F_d(lo) - F_d(hi)) / dx_d, F_1 = D / dx_1 *
//                    (x(i-1) - x(i)), F_2, F_3 = D * (x(j-1) - x(j)) / dx, in the reference's
//                    operation order (this file is built without FMA contraction). The
//                    reference writes the top x3 face flux into the x2 slot
//                    (poisson_equation.hpp:153-157); the operator here is the 7-point one.
//   GatherKernel     bricks of a uniform single-level mesh -> the N^3 level array of the
//   ScatterKernel    multigrid hierarchy (pth_gmg.hip), and back into the interior cells
//
// pth_mg is MGSolver<u, rhs, PoissonEquation> (src/solvers/mg_solver.hpp:37-203) over
// MeshData: setup gathers rhs (scaled onto the unit cube) and the initial u into the solver's
// FP64 level arrays; the solve runs V-cycles in the batched level kernels until the RMS
// residual of the MeshData problem is below the tolerance, scatters u back into its bricks and
// refills u's ghosts with the boundary exchange (the reference's last AddTasks iteration ends
// with the exchange of u, mg_solver.hpp:139-141). The hierarchy is the uniform one of
// src/mesh/mesh-gmg.cpp:103-154 on a uniform mesh (cell counts halving while even), held as
// whole-level arrays: one launch per level stage for every block at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "pth_device.hpp"
#include "pth_gmg.hpp"
#include "pth_internal.hpp"
#include "pth_mgtree.hpp"

namespace pth {
namespace {

constexpr int kWG = 256;
constexpr int kDotMaxWG = 2048;

struct Shape {  // brick shape of a MeshData field
  int e0, e1, e2;     // entire extents
  int g0, g1, g2;     // interior starts
  int n0, n1, n2;     // interior sizes
  long long cells;    // e0 * e1 * e2
};

Shape ShapeOf(const Extent &x) {
  Shape s;
  s.e0 = x.ent(0), s.e1 = x.ent(1), s.e2 = x.ent(2);
  s.g0 = x.g[0], s.g1 = x.g[1], s.g2 = x.g[2];
  s.n0 = x.n[0], s.n1 = x.n[1], s.n2 = x.n[2];
  s.cells = x.cells();
  return s;
}

unsigned Blocks(long long n) { return static_cast<unsigned>((n + kWG - 1) / kWG); }

// brick offset of interior cell t (flattened over [k][j][i] of the interior)
__device__ __forceinline__ long long InteriorCell(const Shape &s, long long t) {
  const int i = static_cast<int>(t % s.n0);
  const long long r = t / s.n0;
  const int j = static_cast<int>(r % s.n1);
  const int k = static_cast<int>(r / s.n1);
  return (static_cast<long long>(k + s.g2) * s.e1 + (j + s.g1)) * s.e0 + (i + s.g0);
}

__global__ void __launch_bounds__(kWG)
    LinearKernel(const double *__restrict__ a, const double *__restrict__ b, double wa,
                 double wb, double *out, long long n) {
  const long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (t >= n) return;
  out[t] = b ? wa * a[t] + wb * b[t] : a[t];
}

// one thread per interior cell of every (block, component) brick
__global__ void __launch_bounds__(kWG) ZeroKernel(double *data, Shape s, long long total) {
  const long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (t >= total) return;
  const long long per = static_cast<long long>(s.n0) * s.n1 * s.n2;
  data[(t / per) * s.cells + InteriorCell(s, t % per)] = 0.0;
}

__device__ __forceinline__ double WorkgroupSum(double v, double *wsum) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kWG / 64; ++w) s += wsum[w];
  return s;
}

// partial[wg] = sum over this workgroup's interior cells (grid-stride over blocks x cells) of
// sum_c a_c * b_c
__global__ void __launch_bounds__(kWG)
    DotKernel(const double *__restrict__ a, const double *__restrict__ b, int ncomp, Shape s,
              long long total, double *__restrict__ partial) {
  __shared__ double wsum[kWG / 64];
  const long long per = static_cast<long long>(s.n0) * s.n1 * s.n2;
  double acc = 0.0;
  for (long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x; t < total;
       t += static_cast<long long>(gridDim.x) * kWG) {
    const long long blk = t / per;
    const long long c = InteriorCell(s, t % per);
    double cell = 0.0;
    for (int v = 0; v < ncomp; ++v) {
      const long long o = (blk * ncomp + v) * s.cells + c;
      cell += a[o] * b[o];
    }
    acc += cell;
  }
  const double sum = WorkgroupSum(acc, wsum);
  if (threadIdx.x == 0) partial[blockIdx.x] = sum;
}

__global__ void __launch_bounds__(kWG)
    SumPartialsKernel(const double *__restrict__ partial, int n, double *__restrict__ out) {
  __shared__ double wsum[kWG / 64];
  double acc = 0.0;
  for (int p = threadIdx.x; p < n; p += kWG) acc += partial[p];
  const double sum = WorkgroupSum(acc, wsum);
  if (threadIdx.x == 0) *out = sum;
}

// Ax on the interior cells of every block; coef[b].dx = the block's cell widths. Df: the face
// field D ([block][3][k][j][i], component d = D on the lower face along d) or null for the
// constant D.
__global__ void __launch_bounds__(kWG)
    PoissonAxKernel(const double *__restrict__ x, double *__restrict__ out, Shape s,
                    const BlockCoefs *__restrict__ coef, double D,
                    const double *__restrict__ Df, double alpha, int ndim, long long total) {
  const long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (t >= total) return;
  const long long per = static_cast<long long>(s.n0) * s.n1 * s.n2;
  const long long blk = t / per;
  const long long cl = InteriorCell(s, t % per);
  const long long c = blk * s.cells + cl;
  const long long p = static_cast<long long>(s.e0) * s.e1;
  // D on the lower and upper face of the cell along each dimension
  const double *d0 = Df ? Df + blk * 3 * s.cells + cl : nullptr;
  const double D1l = Df ? d0[0] : D, D1h = Df ? d0[1] : D;
  const double u = x[c];
  const double dx1 = coef[blk].dx[0];
  double r = -alpha * u;
  const double f1l = D1l / dx1 * (x[c - 1] - u);
  const double f1r = D1h / dx1 * (u - x[c + 1]);
  r += (f1l - f1r) / dx1;
  if (ndim > 1) {
    const double D2l = Df ? d0[s.cells] : D, D2h = Df ? d0[s.cells + s.e0] : D;
    const double dx2 = coef[blk].dx[1];
    const double f2l = D2l * (x[c - s.e0] - u) / dx2;
    const double f2r = D2h * (u - x[c + s.e0]) / dx2;
    r += (f2l - f2r) / dx2;
  }
  if (ndim > 2) {
    const double D3l = Df ? d0[2 * s.cells] : D, D3h = Df ? d0[2 * s.cells + p] : D;
    const double dx3 = coef[blk].dx[2];
    const double f3l = D3l * (x[c - p] - u) / dx3;
    const double f3r = D3h * (u - x[c + p]) / dx3;
    r += (f3l - f3r) / dx3;
  }
  out[c] = r;
}

// level[c] = scale * brick value of global cell c = (K * N + J) * N + I; map: block id of each
// block position (bz * nb1 + by) * nb0 + bx. Consecutive threads take consecutive I, a run
// along x inside one brick row.
__global__ void __launch_bounds__(kWG)
    GatherKernel(const double *__restrict__ brick, const int *__restrict__ map, Shape s, int N,
                 double scale, double *__restrict__ level) {
  const long long c = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (c >= static_cast<long long>(N) * N * N) return;
  const int I = static_cast<int>(c % N), J = static_cast<int>((c / N) % N),
            K = static_cast<int>(c / (static_cast<long long>(N) * N));
  const int nb0 = N / s.n0, nb1 = N / s.n1;
  const int b = map[(K / s.n2 * nb1 + J / s.n1) * nb0 + I / s.n0];
  const long long o = b * s.cells +
                      (static_cast<long long>(K % s.n2 + s.g2) * s.e1 + (J % s.n1 + s.g1)) *
                          s.e0 +
                      (I % s.n0 + s.g0);
  level[c] = scale * brick[o];
}

__global__ void __launch_bounds__(kWG)
    ScatterKernel(const double *__restrict__ level, const int *__restrict__ map, Shape s, int N,
                  double *__restrict__ brick) {
  const long long c = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (c >= static_cast<long long>(N) * N * N) return;
  const int I = static_cast<int>(c % N), J = static_cast<int>((c / N) % N),
            K = static_cast<int>(c / (static_cast<long long>(N) * N));
  const int nb0 = N / s.n0, nb1 = N / s.n1;
  const int b = map[(K / s.n2 * nb1 + J / s.n1) * nb0 + I / s.n0];
  const long long o = b * s.cells +
                      (static_cast<long long>(K % s.n2 + s.g2) * s.e1 + (J % s.n1 + s.g1)) *
                          s.e0 +
                      (I % s.n0 + s.g0);
  brick[o] = level[c];
}

int SolverField(MeshData *md, const char *name, const char *what) {
  const int f = name ? md->mesh->FieldIndex(name) : -1;
  if (f < 0) {
    SetError(std::string(what) + ": no field " + (name ? name : "(null)"));
    return -1;
  }
  return f;
}

pth_status NeedFinalized(const Mesh *m, const char *what) {
  if (!m || !m->finalized) {
    SetError(std::string(what) + ": mesh is not finalized");
    return PTH_FAIL;
  }
  return PTH_COMPLETE;
}

}  // namespace

// ---- the solver handle ------------------------------------------------------------------
struct MgMesh {
  Mesh *mesh = nullptr;
  int fu = -1, frhs = -1;
  int N = 0;
  double scale = 1.0;   // L^2 / D: the MeshData problem on the unit cube
  pth_mg_params P{};
  GmgParams G{};
  std::unique_ptr<GmgSolver1<float>> f;
  std::unique_ptr<GmgSolver1<double>> d;
  int *map = nullptr;   // device: block id by block position
  std::uint64_t generation = ~0ull;
  hipEvent_t ev = nullptr;
  bool ready = false;   // setup ran on the current data
  int iterations = 0;
  double rms = 0.0;
  std::vector<double> history;  // RMS residual before every V-cycle and after the last
  MgTree *tree = nullptr;       // block-mesh hierarchy (meshes other than the uniform cube)
  ~MgMesh() {
    if (tree) MgTreeDestroy(tree);
    if (map) (void)hipFree(map);
    if (ev) (void)hipEventDestroy(ev);
  }
  template <typename F>
  pth_status Do(F fn) {
    return f ? fn(*f) : fn(*d);
  }
  // block map of the current mesh; the mesh must still be uniform, single-level and cubic
  pth_status Map() {
    if (map && generation == mesh->generation) return PTH_COMPLETE;
    const Topology &t = mesh->topo;
    if (t.multilevel) {
      SetError("MGSolver: the mesh is no longer single-level");
      return PTH_FAIL;
    }
    const int nb[3] = {t.nx[0] / t.bnx[0], t.nx[1] / t.bnx[1], t.nx[2] / t.bnx[2]};
    std::vector<int> h(static_cast<size_t>(nb[0]) * nb[1] * nb[2], -1);
    for (size_t b = 0; b < t.blocks.size(); ++b) {
      const Lattice &L = t.blocks[b].at;
      if (L.x[0] < 0 || L.x[0] >= nb[0] || L.x[1] < 0 || L.x[1] >= nb[1] || L.x[2] < 0 ||
          L.x[2] >= nb[2]) {
        SetError("MGSolver: block outside the base lattice");
        return PTH_FAIL;
      }
      h[(L.x[2] * nb[1] + L.x[1]) * nb[0] + L.x[0]] = static_cast<int>(b);
    }
    for (int v : h)
      if (v < 0) {
        SetError("MGSolver: the blocks do not tile the mesh");
        return PTH_FAIL;
      }
    if (map) (void)hipFree(map);
    map = nullptr;
    PTH_HIP(hipMalloc(&map, sizeof(int) * h.size()));
    PTH_HIP(hipMemcpy(map, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    generation = mesh->generation;
    return PTH_COMPLETE;
  }
};

}  // namespace pth

using namespace pth;

struct pth_mg : MgMesh {};

extern "C" {

// ---- solver field utilities ---------------------------------------------------------------
static pth_status Linear(pth_meshdata *mdh, const char *a, const char *b, const char *out,
                         double wa, double wb, const char *what) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (NeedFinalized(m, what) != PTH_COMPLETE) return PTH_FAIL;
  const int fa = SolverField(md, a, what), fo = SolverField(md, out, what);
  const int fb = b ? SolverField(md, b, what) : fa;
  if (fa < 0 || fb < 0 || fo < 0) return PTH_FAIL;
  const int nc = m->fields[fa].ncomp;
  if (m->fields[fb].ncomp != nc || m->fields[fo].ncomp != nc) {
    SetError(std::string(what) + ": fields differ in component count");
    return PTH_FAIL;
  }
  const long long n = md->store[fo].ndata;
  hipLaunchKernelGGL(LinearKernel, dim3(Blocks(n)), dim3(kWG), 0, m->stream,
                     md->store[fa].data, b ? md->store[fb].data : nullptr, wa, wb,
                     md->store[fo].data, n);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status pth_solver_copy_data(pth_meshdata *md, const char *in, const char *out) {
  return Linear(md, in, nullptr, out, 1.0, 0.0, "CopyData");
}

pth_status pth_solver_add_fields_and_store(pth_meshdata *md, const char *a, const char *b,
                                           const char *out, double wa, double wb) {
  if (!b) {
    SetError("AddFieldsAndStore: no field (null)");
    return PTH_FAIL;
  }
  return Linear(md, a, b, out, wa, wb, "AddFieldsAndStore");
}

pth_status pth_solver_set_to_zero(pth_meshdata *mdh, const char *var) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (NeedFinalized(m, "SetToZero") != PTH_COMPLETE) return PTH_FAIL;
  const int f = SolverField(md, var, "SetToZero");
  if (f < 0) return PTH_FAIL;
  const Shape s = ShapeOf(m->topo.fine);
  const long long total = static_cast<long long>(m->topo.blocks.size()) * m->fields[f].ncomp *
                          s.n0 * s.n1 * s.n2;
  if (total == 0) return PTH_COMPLETE;
  hipLaunchKernelGGL(ZeroKernel, dim3(Blocks(total)), dim3(kWG), 0, m->stream,
                     md->store[f].data, s, total);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status pth_solver_dot_product(pth_meshdata *mdh, const char *a, const char *b, double *out) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (!out) {
    SetError("DotProduct: null output");
    return PTH_FAIL;
  }
  if (NeedFinalized(m, "DotProduct") != PTH_COMPLETE) return PTH_FAIL;
  const int fa = SolverField(md, a, "DotProduct"), fb = SolverField(md, b, "DotProduct");
  if (fa < 0 || fb < 0) return PTH_FAIL;
  const int nc = m->fields[fa].ncomp;
  if (m->fields[fb].ncomp != nc) {
    SetError("DotProduct: fields differ in component count");
    return PTH_FAIL;
  }
  const Shape s = ShapeOf(m->topo.fine);
  const long long total =
      static_cast<long long>(m->topo.blocks.size()) * s.n0 * s.n1 * s.n2;
  const int nwg = static_cast<int>(std::max<long long>(
      1, std::min<long long>({static_cast<long long>(Blocks(total)), kDotMaxWG,
                              m->scratch_n - 8})));
  hipLaunchKernelGGL(DotKernel, dim3(nwg), dim3(kWG), 0, m->stream, md->store[fa].data,
                     md->store[fb].data, nc, s, total, m->scratch);
  hipLaunchKernelGGL(SumPartialsKernel, dim3(1), dim3(kWG), 0, m->stream, m->scratch, nwg,
                     m->scratch + m->scratch_n);
  PTH_HIP(hipGetLastError());
  PTH_HIP(hipMemcpyAsync(m->pinned, m->scratch + m->scratch_n, sizeof(double),
                         hipMemcpyDeviceToHost, m->stream));
  PTH_HIP(hipStreamSynchronize(m->stream));
  *out = m->pinned[0];
  return PTH_COMPLETE;
}

// ---- Poisson equation -----------------------------------------------------------------------
pth_status pth_poisson_ax(pth_meshdata *mdh, const pth_poisson_equation *eq, const char *x,
                          const char *out) {
  MeshData *md = Impl(mdh);
  Mesh *m = md->mesh;
  if (!eq) {
    SetError("PoissonEquation::Ax: null equation");
    return PTH_FAIL;
  }
  if (NeedFinalized(m, "PoissonEquation::Ax") != PTH_COMPLETE) return PTH_FAIL;
  const int fx = SolverField(md, x, "PoissonEquation::Ax");
  const int fo = SolverField(md, out, "PoissonEquation::Ax");
  if (fx < 0 || fo < 0) return PTH_FAIL;
  if (m->fields[fx].ncomp != 1 || m->fields[fo].ncomp != 1) {
    SetError("PoissonEquation::Ax: x and out must have one component");
    return PTH_FAIL;
  }
  if (fx == fo) {
    SetError("PoissonEquation::Ax: x and out must be different fields");
    return PTH_FAIL;
  }
  const Topology &t = m->topo;
  for (int d = 0; d < t.ndim; ++d)
    if (t.fine.g[d] < 1) {
      SetError("PoissonEquation::Ax: needs one ghost layer");
      return PTH_FAIL;
    }
  const double *Df = nullptr;
  if (eq->D_field) {
    const int fd = SolverField(md, eq->D_field, "PoissonEquation::Ax");
    if (fd < 0) return PTH_FAIL;
    if (m->fields[fd].ncomp != 3) {
      SetError("PoissonEquation::Ax: the face field D needs 3 components");
      return PTH_FAIL;
    }
    Df = md->store[fd].data;
  }
  const Shape s = ShapeOf(t.fine);
  const long long total = static_cast<long long>(t.blocks.size()) * s.n0 * s.n1 * s.n2;
  if (total == 0) return PTH_COMPLETE;
  hipLaunchKernelGGL(PoissonAxKernel, dim3(Blocks(total)), dim3(kWG), 0, m->stream,
                     md->store[fx].data, md->store[fo].data, s, m->coef, eq->D, Df,
                     eq->diagonal_alpha, t.ndim, total);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

// ---- MGSolver over MeshData -------------------------------------------------------------------
pth_mg *pth_mg_create(pth_mesh *mh, const char *u, const char *rhs, const pth_mg_params *p,
                      const pth_poisson_equation *eq) {
  Mesh *m = Impl(mh);
  if (NeedFinalized(m, "MGSolver") != PTH_COMPLETE) return nullptr;
  if (!p || !eq) {
    SetError("MGSolver: null parameters or equation");
    return nullptr;
  }
  if (p->max_iters < 1 || !(p->residual_tolerance >= 0.0) ||
      (p->precision != 0 && p->precision != 1) || p->inner_cycles < 1 || p->pre_sweeps < 0 ||
      p->post_sweeps < 0 || !(p->omega > 0.0) || p->smoother < 0 || p->smoother > 3) {
    SetError("MGSolver: bad parameters");
    return nullptr;
  }
  if ((!eq->D_field && !(eq->D > 0.0)) || !(eq->diagonal_alpha >= 0.0)) {
    SetError("MGSolver: the equation needs D > 0 and diagonal_alpha >= 0");
    return nullptr;
  }
  int fD = -1;
  if (eq->D_field) {
    fD = m->FieldIndex(eq->D_field);
    if (fD < 0 || m->fields[fD].ncomp != 3) {
      SetError("MGSolver: the face field D must be a 3-component field of the mesh");
      return nullptr;
    }
  }
  const int fu = m->FieldIndex(u ? u : ""), fr = m->FieldIndex(rhs ? rhs : "");
  if (fu < 0 || fr < 0 || fu == fr) {
    SetError("MGSolver: u and rhs must be two fields of the mesh");
    return nullptr;
  }
  if (m->fields[fu].ncomp != 1 || m->fields[fr].ncomp != 1 ||
      !(m->fields[fu].flags & PTH_FILL_GHOST)) {
    SetError("MGSolver: u (FillGhost) and rhs must have one component");
    return nullptr;
  }
  const Topology &t = m->topo;
  const double L = t.xmax[0] - t.xmin[0];
  bool cube = t.ndim == 3 && !t.multilevel && t.nx[0] == t.nx[1] && t.nx[0] == t.nx[2] &&
              t.xmax[1] - t.xmin[1] == L && t.xmax[2] - t.xmin[2] == L &&
              eq->diagonal_alpha == 0.0 && p->smoother == 0 && fD < 0;
  for (int face = 0; face < 6; ++face) cube = cube && t.bc[face] == PTH_BC_FIXED_FACE;
  if (!cube) {
    // block-mesh hierarchy, FP64
    if (p->precision != 1) {
      SetError("MGSolver: mixed precision needs a uniform N^3 cube with fixed_face faces and "
               "the weighted-Jacobi smoother; use precision 1 on this mesh");
      return nullptr;
    }
    for (int face = 0; face < 2 * t.ndim; ++face)
      if (t.bc[face] != PTH_BC_FIXED_FACE && t.bc[face] != PTH_BC_PERIODIC) {
        SetError("MGSolver: every face must be fixed_face or periodic (the poisson_gmg "
                 "package's FixedFace conditions)");
        return nullptr;
      }
    auto *h = new pth_mg();
    h->mesh = m;
    h->fu = fu;
    h->frhs = fr;
    h->P = *p;
    h->tree = MgTreeCreate(m, fu, fr, fD, *p, eq->D, eq->diagonal_alpha);
    if (!h->tree) {
      delete h;
      return nullptr;
    }
    return h;
  }
  auto *h = new pth_mg();
  h->mesh = m;
  h->fu = fu;
  h->frhs = fr;
  h->N = t.nx[0];
  h->scale = L * L / eq->D;
  h->P = *p;
  h->G = {p->precision, p->omega, p->pre_sweeps, p->post_sweeps};
  h->G.inner = p->precision == 0 ? p->inner_cycles : 1;
  pth_status st = PTH_COMPLETE;
  if (p->precision == 0) {
    h->f.reset(new GmgSolver1<float>());
    st = h->f->Init(h->N);
  } else {
    h->d.reset(new GmgSolver1<double>());
    st = h->d->Init(h->N);
  }
  if (st == PTH_COMPLETE && hipEventCreateWithFlags(&h->ev, hipEventDisableTiming) != hipSuccess) {
    SetError("MGSolver: hipEventCreate failed");
    st = PTH_FAIL;
  }
  if (st == PTH_COMPLETE) st = h->Map();
  if (st != PTH_COMPLETE) {
    delete h;
    return nullptr;
  }
  return h;
}

void pth_mg_destroy(pth_mg *h) { delete h; }

pth_status pth_mg_setup(pth_mg *h, pth_meshdata *mdh) {
  MeshData *md = Impl(mdh);
  if (!h || md->mesh != h->mesh) {
    SetError("MGSolver::AddSetupTasks: MeshData of another mesh");
    return PTH_FAIL;
  }
  if (h->tree) {
    h->ready = MgTreeSetup(h->tree, *md) == PTH_COMPLETE;
    return h->ready ? PTH_COMPLETE : PTH_FAIL;
  }
  if (h->Map() != PTH_COMPLETE) return PTH_FAIL;
  Mesh *m = h->mesh;
  const Shape s = ShapeOf(m->topo.fine);
  const long long cells = static_cast<long long>(h->N) * h->N * h->N;
  const pth_status st = h->Do([&](auto &S) -> pth_status {
    hipLaunchKernelGGL(GatherKernel, dim3(Blocks(cells)), dim3(kWG), 0, m->stream,
                       md->store[h->frhs].data, h->map, s, h->N, h->scale, S.p.b64);
    hipLaunchKernelGGL(GatherKernel, dim3(Blocks(cells)), dim3(kWG), 0, m->stream,
                       md->store[h->fu].data, h->map, s, h->N, 1.0, S.p.u[0]);
    PTH_HIP(hipGetLastError());
    S.p.cur = 0;
    PTH_HIP(hipEventRecord(h->ev, m->stream));
    PTH_HIP(hipStreamWaitEvent(S.s, h->ev, 0));
    return S.Norm();
  });
  h->ready = st == PTH_COMPLETE;
  return st;
}

pth_status pth_mg_solve(pth_mg *h, pth_meshdata *mdh) {
  MeshData *md = Impl(mdh);
  if (!h || md->mesh != h->mesh) {
    SetError("MGSolver::AddTasks: MeshData of another mesh");
    return PTH_FAIL;
  }
  if (!h->ready) {
    SetError("MGSolver::AddTasks: the setup has not run");
    return PTH_FAIL;
  }
  if (h->tree) {
    h->ready = false;
    if (MgTreeSolve(h->tree, *md, &h->iterations, &h->history) != PTH_COMPLETE)
      return PTH_FAIL;
    h->rms = h->history.back();
    return pth_exchange_fused(mdh, PTH_BT_ANY);
  }
  if (h->Map() != PTH_COMPLETE) return PTH_FAIL;
  Mesh *m = h->mesh;
  const Shape s = ShapeOf(m->topo.fine);
  const long long cells = static_cast<long long>(h->N) * h->N * h->N;
  // RMS residual of the MeshData problem = ||r'|| / (scale sqrt(cells)) with r' the unit-cube
  // residual; the solver's stop test is relative to ||b'||
  const double rms_of_abs = 1.0 / (h->scale * std::sqrt(static_cast<double>(cells)));
  pth_status st = h->Do([&](auto &S) -> pth_status {
    const double to_abs = S.bnorm > 0.0 ? S.bnorm : 1.0;
    const double tol = h->P.residual_tolerance / (rms_of_abs * to_abs);
    int it = 0;
    double rel = 0.0;
    if (S.Solve(h->G, tol, h->P.max_iters, &it, &rel, 1) != PTH_COMPLETE) return PTH_FAIL;
    h->iterations = it;
    h->history.clear();
    for (double r : S.history) h->history.push_back(r * to_abs * rms_of_abs);
    h->rms = h->history.back();
    PTH_HIP(hipEventRecord(h->ev, S.s));
    PTH_HIP(hipStreamWaitEvent(m->stream, h->ev, 0));
    hipLaunchKernelGGL(ScatterKernel, dim3(Blocks(cells)), dim3(kWG), 0, m->stream,
                       S.p.u[S.p.cur], h->map, s, h->N, md->store[h->fu].data);
    PTH_HIP(hipGetLastError());
    return PTH_COMPLETE;
  });
  h->ready = false;
  if (st != PTH_COMPLETE) return PTH_FAIL;
  return pth_exchange_fused(mdh, PTH_BT_ANY);
}

pth_status pth_mg_result(const pth_mg *h, int *iterations, double *rms_residual) {
  if (!h) {
    SetError("MGSolver: null solver");
    return PTH_FAIL;
  }
  if (iterations) *iterations = h->iterations;
  if (rms_residual) *rms_residual = h->rms;
  return PTH_COMPLETE;
}

int pth_mg_levels(const pth_mg *h) {
  if (!h) return 0;
  if (h->tree) return MgTreeLevels(h->tree);
  return h->f ? static_cast<int>(h->f->p.sizes.size()) : static_cast<int>(h->d->p.sizes.size());
}

pth_status pth_mg_set_graph(pth_mg *h, int on) {
  if (!h) {
    SetError("MGSolver: null solver");
    return PTH_FAIL;
  }
  if (h->tree) MgTreeSetGraph(h->tree, on != 0);
  return PTH_COMPLETE;
}

int pth_mg_history(const pth_mg *h, double *out, int max) {
  if (!h) return 0;
  const int n = std::min(static_cast<int>(h->history.size()), std::max(max, 0));
  for (int q = 0; q < n; ++q) out[q] = h->history[q];
  return static_cast<int>(h->history.size());
}

}  // extern "C"
