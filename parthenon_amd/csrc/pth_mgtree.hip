// Multigrid Poisson solver over MeshData on meshes of any static refinement, any dimension
// and any mix of FixedFace / periodic faces: the general form of MGSolver<u, rhs,
// PoissonEquation> (src/solvers/mg_solver.hpp:73-520) on a hierarchy of block meshes.
//
// Hierarchy (the levels of src/mesh/mesh-gmg.cpp:103-154, finest first):
//   * the leaf mesh itself;
//   * the mesh capped at each coarser refinement level down to the base grid: every block of
//     the capped level whose children exist becomes one block (its children restrict into
//     it, the reference's gmg_restrict_send to gmg_coarser_neighbors), every other block is
//     the same block on both levels (the reference shares those MeshBlocks between its grids:
//     gmg_leaf_neighbors, same_to_same buffers, bvals_utils.hpp:49-63);
//   * below the base grid, the tree levels of the forest (2^d base blocks merge into one, as
//     many times as the forest's root level, forest.cpp:64-105), then levels whose blocks keep
//     their positions and halve their cells (the reference's negative logical levels) while
//     every block size is even, down to one cell per block (one ghost layer on these levels)
//     and two cells per dimension for the level (where the reference's single-tree forest
//     ends with one cell).
// Every level is a full tiling of the domain, so each level's boundary exchange is the leaf
// exchange of that level's mesh (same-level copies, restriction into coarse buffers and
// linear prolongation at fine-coarse faces, FixedFace conditions). The reference's grids
// instead hold only the level's blocks and the coarser leaves beside them
// (two_level_composite, tree.cpp:163-230) and smooth only those; here the unrefined part of
// each level is smoothed as well. The fixed point is the same: the solve stops on the RMS
// residual of the leaf problem, measured with the leaf exchange and PoissonEquation::Ax as in
// AddTasks (mg_solver.hpp:131-158).
//
// Cycle (correction scheme; the reference cycles the full solution, FAS, which gives the same
// iterates for this linear operator in exact arithmetic): on each level, pre-smoothing,
// residual r = b - A x, restriction of r (RestrictAverage) into the coarser level's b with a
// zero initial x there, the coarser cycle, prolongation of its x (ProlongateSharedLinear from
// the coarse cells and their face neighbours, the reference's gmg_prolongate_recv with
// ProResInfo::GetInteriorProlongate, bnd_info.cpp:371-385) added to x, post-smoothing; the
// coarsest level only pre-smooths (mg_solver.hpp:441-448). The smoother is the reference's
// SRJ (stages 1-3 with the Yang & Mittal weights of mg_solver.hpp:312-345), or weighted Jacobi
// with explicit weight and sweep counts. One Jacobi stage (mg_solver.hpp:222-297, 299-311):
// exchange x, t = w (b - (A x - diag x)) / diag + (1 - w) x with the unigrid diagonal of
// PoissonEquation::SetDiagonal (poisson_equation.hpp:86-125), x = t.
//
// Transfers between levels are region tables walked by the exchange's persistent restriction,
// prolongation and box-copy kernels. FP64 throughout.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "pth_burgers.hpp"
#include "pth_internal.hpp"
#include "pth_mgtree.hpp"

namespace pth {
namespace {

constexpr int kWG = 256;
enum { kX = 0, kT = 1, kB = 2, kR = 3, kD = 4 };  // level fields: solution/correction,
                                                  // Jacobi output, right-hand side,
                                                  // residual, face field D (3 components)

struct BrickShape {
  int e0, e1;
  int g0, g1, g2;
  int n0, n1, n2;
  long long cells;
};

BrickShape ShapeOf(const Extent &x) {
  BrickShape s;
  s.e0 = x.ent(0), s.e1 = x.ent(1);
  s.g0 = x.g[0], s.g1 = x.g[1], s.g2 = x.g[2];
  s.n0 = x.n[0], s.n1 = x.n[1], s.n2 = x.n[2];
  s.cells = x.cells();
  return s;
}

unsigned Grid(long long n) { return static_cast<unsigned>((n + kWG - 1) / kWG); }

// brick offset of block `blk`'s interior cell t (flattened over its [k][j][i] interior)
__device__ __forceinline__ long long Cell(const BrickShape &s, long long t) {
  const long long per = static_cast<long long>(s.n0) * s.n1 * s.n2;
  const long long blk = t / per;
  const long long r = t % per;
  const int i = static_cast<int>(r % s.n0);
  const long long q = r / s.n0;
  const int j = static_cast<int>(q % s.n1);
  const int k = static_cast<int>(q / s.n1);
  return blk * s.cells + (static_cast<long long>(k + s.g2) * s.e1 + (j + s.g1)) * s.e0 +
         (i + s.g0);
}

// MODE 0: Jacobi stage out = w (b - (A x - diag x)) / diag + (1 - w) x
// MODE 1: residual out = 1.0 b + (-1.0) A x (AddFieldsAndStoreInteriorSelect<rhs, Ax, res>)
// A x in the operation order of CalculateFluxes + FluxMultiplyMatrix, diag as SetDiagonal;
// D on the cell's faces from the face field Df ([block][3] bricks, component d = D on the
// lower face along d) or the constant D when Df is null.
template <int MODE>
__global__ void __launch_bounds__(kWG)
    LevelCellKernel(const double *__restrict__ x, const double *__restrict__ b,
                    double *__restrict__ out, BrickShape s, const BlockCoefs *__restrict__ coef,
                    double D, const double *__restrict__ Df, double alpha, double w, int ndim,
                    long long total) {
  const long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (t >= total) return;
  const long long blk = t / (static_cast<long long>(s.n0) * s.n1 * s.n2);
  const long long c = Cell(s, t);
  const long long p = static_cast<long long>(s.e0) * s.e1;
  const double *d0 = Df ? Df + blk * 3 * s.cells + (c - blk * s.cells) : nullptr;  // comp 0
  const double u = x[c];
  const double dx1 = coef[blk].dx[0];
  const double D1l = Df ? d0[0] : D, D1h = Df ? d0[1] : D;
  double ax = -alpha * u;
  const double f1l = D1l / dx1 * (x[c - 1] - u);
  const double f1r = D1h / dx1 * (u - x[c + 1]);
  ax += (f1l - f1r) / dx1;
  double diag = -(D1l + D1h) / (dx1 * dx1) - alpha;
  if (ndim > 1) {
    const double D2l = Df ? d0[s.cells] : D, D2h = Df ? d0[s.cells + s.e0] : D;
    const double dx2 = coef[blk].dx[1];
    const double f2l = D2l * (x[c - s.e0] - u) / dx2;
    const double f2r = D2h * (u - x[c + s.e0]) / dx2;
    ax += (f2l - f2r) / dx2;
    diag -= (D2l + D2h) / (dx2 * dx2);
  }
  if (ndim > 2) {
    const double D3l = Df ? d0[2 * s.cells] : D, D3h = Df ? d0[2 * s.cells + p] : D;
    const double dx3 = coef[blk].dx[2];
    const double f3l = D3l * (x[c - p] - u) / dx3;
    const double f3r = D3h * (u - x[c + p]) / dx3;
    ax += (f3l - f3r) / dx3;
    diag -= (D3l + D3h) / (dx3 * dx3);
  }
  if (MODE == 0) {
    const double off_diag = ax - diag * u;
    const double val = b[c] - off_diag;
    out[c] = w * (val / diag) + (1.0 - w) * u;
  } else {
    out[c] = 1.0 * b[c] + (-1.0) * ax;
  }
}

// Face-field restriction (RestrictAverage on faces) from a block into its coarser-level
// counterpart: coarse face (I, J, K) along direction d, with I in [0, nc_d] and the transverse
// indices in [0, nc), is the mean of the fine faces it covers (2 x 2 in 3-D, 2 in 2-D, one in
// 1-D), fine face index 2 I along d.
struct FaceXfer {
  const double *f;  // the fine block's D brick, component 0
  double *c;        // the coarse block's D brick, component 0
  int o[3];         // octant of the fine block in the coarse one
  int pad;
};
struct FaceGeom {
  int fe0, fe1, ce0, ce1;
  long long fcells, ccells;
  int fg[3], cg[3], nc[3], act[3];
  int ndim;
};

__global__ void __launch_bounds__(kWG)
    FaceRestrictKernel(const FaceXfer *__restrict__ recs, FaceGeom G, long long per_rec,
                       long long total) {
  const long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (t >= total) return;
  const FaceXfer R = recs[t / per_rec];
  long long q = t % per_rec;
  int P[3];
  for (int e = 0; e < 3; ++e) P[e] = G.act[e] ? G.nc[e] + 1 : 1;
  const long long per_d = static_cast<long long>(P[0]) * P[1] * P[2];
  const int d = static_cast<int>(q / per_d);
  q %= per_d;
  int I[3];
  I[0] = static_cast<int>(q % P[0]);
  I[1] = static_cast<int>((q / P[0]) % P[1]);
  I[2] = static_cast<int>(q / (static_cast<long long>(P[0]) * P[1]));
  for (int e = 0; e < 3; ++e)
    if (e != d && G.act[e] && I[e] >= G.nc[e]) return;  // padding of a transverse range
  int f[3], n2[3];
  for (int e = 0; e < 3; ++e) {
    f[e] = G.act[e] ? G.fg[e] + 2 * I[e] : 0;
    n2[e] = (e != d && G.act[e]) ? 2 : 1;
  }
  double sum = 0.0;
  for (int kk = 0; kk < n2[2]; ++kk)
    for (int jj = 0; jj < n2[1]; ++jj)
      for (int ii = 0; ii < n2[0]; ++ii)
        sum += R.f[d * G.fcells + (static_cast<long long>(f[2] + kk) * G.fe1 + (f[1] + jj)) *
                                      G.fe0 +
                   (f[0] + ii)];
  const int cnt = n2[0] * n2[1] * n2[2];
  int C[3];
  for (int e = 0; e < 3; ++e) C[e] = G.act[e] ? G.cg[e] + R.o[e] * G.nc[e] + I[e] : 0;
  R.c[d * G.ccells + (static_cast<long long>(C[2]) * G.ce1 + C[1]) * G.ce0 + C[0]] =
      sum / cnt;
}

// interior cells of every block: MODE 0: a = 0; MODE 1: a += b
template <int MODE>
__global__ void __launch_bounds__(kWG)
    InteriorKernel(double *__restrict__ a, const double *__restrict__ b, BrickShape s,
                   long long total) {
  const long long t = static_cast<long long>(blockIdx.x) * kWG + threadIdx.x;
  if (t >= total) return;
  const long long c = Cell(s, t);
  if (MODE == 0)
    a[c] = 0.0;
  else
    a[c] = a[c] + b[c];
}

// SRJ weights of mg_solver.hpp:312-345 (Yang & Mittal 2017), by dimension and stage count
constexpr double kSrj2[3][2] = {{0.8723, 0.5395}, {1.3895, 0.5617}, {1.7319, 0.5695}};
constexpr double kSrj3[3][3] = {
    {0.9372, 0.6667, 0.5173}, {1.6653, 0.8000, 0.5264}, {2.2473, 0.8571, 0.5296}};

std::int64_t Offset(const Extent &e, int i, int j, int k) {
  return (static_cast<std::int64_t>(k) * e.ent(1) + j) * e.ent(0) + i;
}

Box Interior(const Extent &e) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = e.g[d];
    b.hi[d] = e.g[d] + e.n[d] - 1;
  }
  return b;
}

Box Whole(const Extent &e) {
  Box b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = 0;
    b.hi[d] = e.ent(d) - 1;
  }
  return b;
}

Region Rec(double *a, double *b, const Box &box, int gid, const Extent &ea, const Extent *eb,
           int nc = 1) {
  Region r{};
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

Lattice UpOne(const Lattice &a, const bool *active) {
  Lattice p;
  p.level = a.level - 1;
  for (int d = 0; d < 3; ++d) p.x[d] = active[d] ? a.x[d] / 2 : a.x[d];
  return p;
}

}  // namespace

// How the blocks of a level relate to the next coarser level's.
enum class Coarsen {
  kCap,    // blocks of the finest refinement level restrict into their parents; others are
           // the same blocks
  kMerge,  // every base block restricts into its parent of the next tree level
  kShrink  // every block keeps its position and halves its cells
};

struct TreeLevel {
  std::unique_ptr<Mesh> mesh;
  MeshData *md = nullptr;
  Coarsen to_coarser = Coarsen::kCap;
  // to the next coarser level: r -> r's coarse buffers (restrict), then boxes -> its b (copy);
  // its x -> t's coarse buffers or t (copy), then t's coarse buffers -> t (prolong)
  Table restrict_r, copy_down, copy_up, prolong_t;
  // the face field D to the next coarser level: same blocks copied, the others face-averaged
  Table copy_D;
  std::vector<FaceXfer> face_host;
  FaceXfer *face = nullptr;
  FaceGeom face_geom{};
  long long cells = 0;  // interior cells of all blocks
  ~TreeLevel() {
    restrict_r.Free();
    copy_down.Free();
    copy_up.Free();
    prolong_t.Free();
    copy_D.Free();
    if (face) (void)hipFree(face);
  }
};

class MgTree {
 public:
  Mesh *user = nullptr;
  int fu = -1, frhs = -1, fD = -1;
  pth_mg_params P{};
  double D = 1.0, alpha = 0.0;
  std::vector<std::unique_ptr<TreeLevel>> lv;  // finest (a copy of the user's mesh) first
  std::vector<double> w_pre, w_post;           // Jacobi weight of each smoothing stage
  // a V-cycle is ~10 launches per level, each a few microseconds of work on the coarse
  // levels: after one eager V-cycle (which builds every level's exchange tables) the next is
  // captured and later V-cycles replay it
  bool use_graph = true, warmed = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipStream_t capture_stream = nullptr;

  ~MgTree() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
  }
  pth_status Build();
  pth_status Setup(MeshData &md);
  pth_status Solve(MeshData &md, int *iters, std::vector<double> *history);

 private:
  pth_status VCycle();
  pth_status NewLevel(const Topology &t, bool coarse_buffers, int x_prolong);
  pth_status BuildTransfers(int l);
  pth_status Exchange(int l, int field);
  pth_status Cells(int l, int mode, int in, int out, double w);
  pth_status Smooth(int l, const std::vector<double> &w);
  pth_status Cycle(int l);
  pth_status Rms(double *rms);
  double *Data(int l, int f) { return lv[l]->md->store[f].data; }
};

pth_status MgTree::NewLevel(const Topology &t, bool coarse_buffers, int x_prolong) {
  std::unique_ptr<TreeLevel> L(new TreeLevel());
  L->mesh.reset(new Mesh());
  Mesh &m = *L->mesh;
  m.topo = t;
  m.force_coarse = coarse_buffers;
  m.stream = user->stream;
  const unsigned fill = PTH_FILL_GHOST | PTH_INDEPENDENT;
  m.fields.push_back(FieldSpec{"x", 1, fill, x_prolong});
  m.fields.push_back(FieldSpec{"t", 1, 0u, PTH_PROLONG_LINEAR});
  m.fields.push_back(FieldSpec{"b", 1, 0u, PTH_PROLONG_LINEAR});
  m.fields.push_back(FieldSpec{"r", 1, 0u, PTH_PROLONG_LINEAR});
  m.fields.push_back(FieldSpec{"D", 3, 0u, PTH_PROLONG_LINEAR});
  if (m.Finalize() != PTH_COMPLETE) return PTH_FAIL;
  L->md = m.Get("base");
  if (!L->md) return PTH_FAIL;
  const Extent &e = m.topo.fine;
  L->cells = static_cast<long long>(m.topo.blocks.size()) * e.n[0] * e.n[1] * e.n[2];
  lv.push_back(std::move(L));
  return PTH_COMPLETE;
}

pth_status MgTree::Build() {
  const Topology &ut = user->topo;
  // the level weights
  const int nd = std::max(ut.ndim, 1) - 1;
  if (P.smoother == 0) {
    w_pre.assign(P.pre_sweeps, P.omega);
    w_post.assign(P.post_sweeps, P.omega);
  } else {
    std::vector<double> w;
    if (P.smoother == 1) w = {1.0};
    if (P.smoother == 2) w = {kSrj2[nd][0], kSrj2[nd][1]};
    if (P.smoother == 3) w = {kSrj3[nd][0], kSrj3[nd][1], kSrj3[nd][2]};
    w_pre = w_post = w;
  }
  int top = 0;
  for (const auto &a : ut.leaf) top = std::max(top, a.level);
  // refined levels: the leaf mesh capped at level L = top .. 0
  for (int L = top; L >= 0; --L) {
    Topology t = ut;
    t.leaf.clear();
    t.inner.clear();
    for (const auto &a : ut.leaf)
      if (a.level <= L) t.leaf.insert(a);
    for (const auto &a : ut.inner) {
      if (a.level < L) t.inner.insert(a);
      if (a.level == L) t.leaf.insert(a);
    }
    t.multilevel = !t.inner.empty();
    const int xp = L == top ? user->fields[fu].prolong : PTH_PROLONG_LINEAR;
    if (NewLevel(t, true, xp) != PTH_COMPLETE) return PTH_FAIL;
    lv.back()->to_coarser = Coarsen::kCap;
  }
  // below the base grid: the forest's tree levels (root_level merges of 2^d base blocks,
  // forest.cpp:64-105), then levels that halve every block's cells as many times as the block
  // size has trailing binary zeros (mesh-gmg.cpp:106-118), with one ghost layer
  int nx[3], bnx[3], ng = ut.ng;
  for (int d = 0; d < 3; ++d) {
    nx[d] = ut.nx[d];
    bnx[d] = ut.bnx[d];
  }
  int merges = ut.RootLevel(), shrinks = 1 << 30;
  for (int d = 0; d < 3; ++d) {
    if (!ut.active[d]) continue;
    int z = 0;
    while (((bnx[d] >> z) & 1) == 0) ++z;
    shrinks = std::min(shrinks, z);
  }
  if (shrinks == (1 << 30)) shrinks = 0;
  for (int q = 0; q < merges + shrinks; ++q) {
    const bool merge = q < merges;
    // a level keeps at least two cells along every active dimension (a mesh of one cell
    // along a dimension is a mesh without that dimension here)
    bool fits = true;
    for (int d = 0; d < 3; ++d) fits = fits && (!ut.active[d] || nx[d] / 2 >= 2);
    if (!fits) break;
    lv.back()->to_coarser = merge ? Coarsen::kMerge : Coarsen::kShrink;
    for (int d = 0; d < 3; ++d) {
      if (!ut.active[d]) continue;
      nx[d] /= 2;
      if (!merge) bnx[d] /= 2;
    }
    if (!merge) ng = 1;
    Topology t;
    const int per[3] = {ut.periodic[0], ut.periodic[1], ut.periodic[2]};
    const std::string err = t.Init(nx, bnx, ng, ut.xmin, ut.xmax, per);
    if (!err.empty()) {
      SetError("MGSolver hierarchy: " + err);
      return PTH_FAIL;
    }
    for (int f = 0; f < 6; ++f) t.bc[f] = ut.bc[f];
    if (NewLevel(t, true, PTH_PROLONG_LINEAR) != PTH_COMPLETE) return PTH_FAIL;
  }
  for (int l = 0; l + 1 < static_cast<int>(lv.size()); ++l)
    if (BuildTransfers(l) != PTH_COMPLETE) return PTH_FAIL;
  return PTH_COMPLETE;
}

pth_status MgTree::BuildTransfers(int l) {
  TreeLevel &F = *lv[l], &C = *lv[l + 1];
  const Topology &ft = F.mesh->topo, &ct = C.mesh->topo;
  const Extent &fe = ft.fine, &ce = ft.coarse, &Ce = ct.fine;
  int top = 0;
  for (const auto &a : ft.leaf) top = std::max(top, a.level);
  FieldStore &r = F.md->store[kR], &t = F.md->store[kT];
  FieldStore &cb = C.md->store[kB], &cx = C.md->store[kX];
  for (Table *q : {&F.restrict_r, &F.prolong_t}) q->item_elems = kProResItemElems;
  for (int fb = 0; fb < static_cast<int>(ft.blocks.size()); ++fb) {
    const Lattice &a = ft.blocks[fb].at;
    bool child = false;
    Lattice target = a;
    int o[3] = {0, 0, 0};
    if (F.to_coarser == Coarsen::kCap) {
      child = a.level == top && a.level > 0;
      if (child) target = UpOne(a, ft.active);
    } else if (F.to_coarser == Coarsen::kMerge) {
      child = true;
      target = a;
      for (int d = 0; d < 3; ++d) target.x[d] = ft.active[d] ? a.x[d] / 2 : a.x[d];
    }
    if (child)
      for (int d = 0; d < 3; ++d) o[d] = ft.active[d] ? static_cast<int>(a.x[d] - 2 * target.x[d]) : 0;
    auto it = ct.id.find(target);
    if (it == ct.id.end()) {
      SetError("MGSolver hierarchy: a block has no counterpart on the coarser level");
      return PTH_FAIL;
    }
    const int cbk = it->second;
    double *r_f = r.data + fb * fe.cells(), *r_c = r.coarse + fb * ce.cells();
    double *t_f = t.data + fb * fe.cells(), *t_c = t.coarse + fb * ce.cells();
    double *b_C = cb.data + cbk * Ce.cells(), *x_C = cx.data + cbk * Ce.cells();
    if (F.to_coarser == Coarsen::kCap && !child) {  // the same block on both levels
      if (fe.n[0] != Ce.n[0] || fe.n[1] != Ce.n[1] || fe.n[2] != Ce.n[2]) {
        SetError("MGSolver hierarchy: block shapes differ between levels");
        return PTH_FAIL;
      }
      F.copy_down.host.push_back(
          Rec(r_f, b_C + Offset(Ce, Ce.g[0], Ce.g[1], Ce.g[2]), Interior(fe), fb, fe, &Ce));
      F.copy_up.host.push_back(
          Rec(x_C, t_f + Offset(fe, fe.g[0], fe.g[1], fe.g[2]), Interior(Ce), cbk, Ce, &fe));
      F.copy_D.host.push_back(Rec(F.md->store[kD].data + fb * 3 * fe.cells(),
                                  C.md->store[kD].data + cbk * 3 * Ce.cells(), Whole(fe), fb,
                                  fe, &Ce, 3));
      continue;
    }
    F.face_host.push_back(FaceXfer{F.md->store[kD].data + fb * 3 * fe.cells(),
                                   C.md->store[kD].data + cbk * 3 * Ce.cells(),
                                   {o[0], o[1], o[2]}, 0});
    // restriction of the block's interior into its coarse buffer, then into the target
    const Box ci = Interior(ce);
    F.restrict_r.host.push_back(Rec(r_f, r_c, ci, fb, fe, &ce));
    int dst[3], src_lo[3];
    for (int d = 0; d < 3; ++d) {
      dst[d] = Ce.g[d] + o[d] * ce.n[d];
      src_lo[d] = dst[d] - ce.g[d];
    }
    F.copy_down.host.push_back(Rec(r_c, b_C + Offset(Ce, dst[0], dst[1], dst[2]), ci, fb, ce, &Ce));
    // the target's cells over the block and a ghost band of m >= 1 cells (the prolongation
    // stencil reads one) -> t's coarse buffer
    Box sb;
    for (int d = 0; d < 3; ++d) {
      const int m = std::min(ce.g[d], Ce.g[d]);
      sb.lo[d] = src_lo[d] + ce.g[d] - m;
      sb.hi[d] = sb.lo[d] + ce.n[d] + 2 * m - 1;
      if (sb.lo[d] < 0 || sb.hi[d] >= Ce.ent(d) || (ft.active[d] && m < 1)) {
        SetError("MGSolver hierarchy: prolongation source outside the coarser block (level " +
                 std::to_string(l) + ", block " + std::to_string(fb) + ", dimension " +
                 std::to_string(d) + ": cells " + std::to_string(sb.lo[d]) + ".." +
                 std::to_string(sb.hi[d]) + " of " + std::to_string(Ce.ent(d)) + ")");
        return PTH_FAIL;
      }
    }
    const int m0 = std::min(ce.g[0], Ce.g[0]), m1 = std::min(ce.g[1], Ce.g[1]),
              m2 = std::min(ce.g[2], Ce.g[2]);
    F.copy_up.host.push_back(
        Rec(x_C, t_c + Offset(ce, ce.g[0] - m0, ce.g[1] - m1, ce.g[2] - m2), sb, cbk, Ce, &ce));
    F.prolong_t.host.push_back(Rec(t_f, t_c, ci, fb, fe, &ce));
  }
  for (Table *q : {&F.restrict_r, &F.copy_down, &F.copy_up, &F.prolong_t, &F.copy_D})
    if (q->Upload() != PTH_COMPLETE) return PTH_FAIL;
  FaceGeom &G = F.face_geom;
  G.fe0 = fe.ent(0), G.fe1 = fe.ent(1), G.ce0 = Ce.ent(0), G.ce1 = Ce.ent(1);
  G.fcells = fe.cells(), G.ccells = Ce.cells();
  for (int d = 0; d < 3; ++d) {
    G.fg[d] = fe.g[d], G.cg[d] = Ce.g[d], G.nc[d] = ce.n[d];
    G.act[d] = ft.active[d] ? 1 : 0;
    // the upper face of the last cell lies in the first ghost cell on both levels
    if (ft.active[d] && (fe.g[d] < 1 || Ce.g[d] < 1)) {
      SetError("MGSolver hierarchy: the face field needs one ghost layer");
      return PTH_FAIL;
    }
  }
  G.ndim = ft.ndim;
  if (!F.face_host.empty()) {
    PTH_HIP(hipMalloc(&F.face, sizeof(FaceXfer) * F.face_host.size()));
    PTH_HIP(hipMemcpy(F.face, F.face_host.data(), sizeof(FaceXfer) * F.face_host.size(),
                      hipMemcpyHostToDevice));
  }
  return PTH_COMPLETE;
}

pth_status MgTree::Exchange(int l, int field) {
  (void)field;  // x is the only FillGhost field of a level
  return pth_exchange_fused(reinterpret_cast<pth_meshdata *>(lv[l]->md), PTH_BT_ANY);
}

pth_status MgTree::Cells(int l, int mode, int in, int out, double w) {
  Mesh &m = *lv[l]->mesh;
  const long long total = lv[l]->cells;
  const BrickShape s = ShapeOf(m.topo.fine);
  const double *Df = fD >= 0 ? Data(l, kD) : nullptr;
  ProfScope prof(mode == 0 ? "mg_level_jacobi" : "mg_level_residual", m.stream);
  if (mode == 0)
    hipLaunchKernelGGL(LevelCellKernel<0>, dim3(Grid(total)), dim3(kWG), 0, m.stream,
                       Data(l, in), Data(l, kB), Data(l, out), s, m.coef, D, Df, alpha, w,
                       m.topo.ndim, total);
  else
    hipLaunchKernelGGL(LevelCellKernel<1>, dim3(Grid(total)), dim3(kWG), 0, m.stream,
                       Data(l, in), Data(l, kB), Data(l, out), s, m.coef, D, Df, alpha, w,
                       m.topo.ndim, total);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status MgTree::Smooth(int l, const std::vector<double> &w) {
  Mesh &m = *lv[l]->mesh;
  for (double wk : w) {
    if (Exchange(l, kX) != PTH_COMPLETE) return PTH_FAIL;
    if (Cells(l, 0, kX, kT, wk) != PTH_COMPLETE) return PTH_FAIL;
    PTH_HIP(hipMemcpyAsync(Data(l, kX), Data(l, kT), sizeof(double) * lv[l]->md->store[kX].ndata,
                           hipMemcpyDeviceToDevice, m.stream));
  }
  return PTH_COMPLETE;
}

pth_status MgTree::Cycle(int l) {
  const int last = static_cast<int>(lv.size()) - 1;
  if (Smooth(l, w_pre) != PTH_COMPLETE) return PTH_FAIL;
  if (l == last) return PTH_COMPLETE;
  TreeLevel &F = *lv[l], &C = *lv[l + 1];
  Mesh &fm = *F.mesh, &cm = *C.mesh;
  // residual, restricted into the coarser level's right-hand side; zero initial correction
  if (Exchange(l, kX) != PTH_COMPLETE) return PTH_FAIL;
  if (Cells(l, 1, kX, kR, 0.0) != PTH_COMPLETE) return PTH_FAIL;
  if (LaunchRestrictT(F.restrict_r, fm.topo, fm.geom, fm.stream) != PTH_COMPLETE) return PTH_FAIL;
  if (LaunchCopy(F.copy_down, fm.stream) != PTH_COMPLETE) return PTH_FAIL;
  const BrickShape cs = ShapeOf(cm.topo.fine);
  hipLaunchKernelGGL(InteriorKernel<0>, dim3(Grid(C.cells)), dim3(kWG), 0, cm.stream,
                     Data(l + 1, kX), nullptr, cs, C.cells);
  PTH_HIP(hipGetLastError());
  if (Cycle(l + 1) != PTH_COMPLETE) return PTH_FAIL;
  // the coarser correction, prolongated and added
  if (Exchange(l + 1, kX) != PTH_COMPLETE) return PTH_FAIL;
  if (LaunchCopy(F.copy_up, fm.stream) != PTH_COMPLETE) return PTH_FAIL;
  if (LaunchProlongT(F.prolong_t, fm.topo, fm.geom, fm.stream, false) != PTH_COMPLETE)
    return PTH_FAIL;
  const BrickShape fs = ShapeOf(fm.topo.fine);
  hipLaunchKernelGGL(InteriorKernel<1>, dim3(Grid(F.cells)), dim3(kWG), 0, fm.stream,
                     Data(l, kX), Data(l, kT), fs, F.cells);
  PTH_HIP(hipGetLastError());
  return Smooth(l, w_post);
}

pth_status MgTree::VCycle() {
  if (!use_graph) return Cycle(0);
  if (!warmed) {
    warmed = true;
    return Cycle(0);
  }
  hipStream_t s = lv[0]->mesh->stream;
  if (!graph_exec) {
    // the caller's stream may be the null stream, which cannot be captured: the levels launch
    // on a stream of the solver's own while the V-cycle is captured
    if (!capture_stream)
      PTH_HIP(hipStreamCreateWithFlags(&capture_stream, hipStreamNonBlocking));
    for (auto &L : lv) L->mesh->stream = capture_stream;
    const bool prof = Prof().on;
    Prof().on = false;  // no event records inside the captured V-cycle
    hipError_t e = hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal);
    const bool began = e == hipSuccess;
    pth_status st = PTH_FAIL;
    hipGraph_t g = nullptr;
    if (began) {
      st = Cycle(0);
      e = hipStreamEndCapture(capture_stream, &g);
    }
    Prof().on = prof;
    for (auto &L : lv) L->mesh->stream = s;
    if (e != hipSuccess && (!began || st == PTH_COMPLETE))
      SetError(std::string("MGSolver: V-cycle capture: ") + hipGetErrorString(e));
    if (st != PTH_COMPLETE || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return PTH_FAIL;
    }
    graph = g;
    PTH_HIP(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0));
  }
  ProfScope prof("mg_vcycle_graph", s);
  PTH_HIP(hipGraphLaunch(graph_exec, s));
  return PTH_COMPLETE;
}

pth_status MgTree::Rms(double *rms) {
  if (Exchange(0, kX) != PTH_COMPLETE) return PTH_FAIL;
  if (Cells(0, 1, kX, kR, 0.0) != PTH_COMPLETE) return PTH_FAIL;
  double sum = 0.0;
  if (pth_solver_dot_product(reinterpret_cast<pth_meshdata *>(lv[0]->md), "r", "r", &sum) !=
      PTH_COMPLETE)
    return PTH_FAIL;
  *rms = std::sqrt(sum / static_cast<double>(lv[0]->cells));
  return PTH_COMPLETE;
}

pth_status MgTree::Setup(MeshData &md) {
  // the leaf level is a copy of the user's mesh: same blocks in the same order
  Mesh &m0 = *lv[0]->mesh;
  const Topology &ut = user->topo;
  if (ut.blocks.size() != m0.topo.blocks.size()) {
    SetError("MGSolver: the mesh changed since the solver was created");
    return PTH_FAIL;
  }
  for (size_t b = 0; b < ut.blocks.size(); ++b)
    if (!(ut.blocks[b].at == m0.topo.blocks[b].at)) {
      SetError("MGSolver: the mesh changed since the solver was created");
      return PTH_FAIL;
    }
  const size_t n = sizeof(double) * md.store[fu].ndata;
  PTH_HIP(hipMemcpyAsync(Data(0, kX), md.store[fu].data, n, hipMemcpyDeviceToDevice, m0.stream));
  PTH_HIP(hipMemcpyAsync(Data(0, kB), md.store[frhs].data, n, hipMemcpyDeviceToDevice,
                         m0.stream));
  if (fD < 0) return PTH_COMPLETE;
  // AddSetupTasks: the face field D (GMGRestrict) down the levels
  PTH_HIP(hipMemcpyAsync(Data(0, kD), md.store[fD].data, sizeof(double) * md.store[fD].ndata,
                         hipMemcpyDeviceToDevice, m0.stream));
  for (int l = 0; l + 1 < static_cast<int>(lv.size()); ++l) {
    TreeLevel &F = *lv[l];
    if (LaunchCopy(F.copy_D, F.mesh->stream) != PTH_COMPLETE) return PTH_FAIL;
    if (F.face_host.empty()) continue;
    const FaceGeom &G = F.face_geom;
    long long per_rec = G.ndim;
    for (int d = 0; d < 3; ++d) per_rec *= G.act[d] ? G.nc[d] + 1 : 1;
    const long long total = per_rec * static_cast<long long>(F.face_host.size());
    hipLaunchKernelGGL(FaceRestrictKernel, dim3(Grid(total)), dim3(kWG), 0, F.mesh->stream,
                       F.face, G, per_rec, total);
    PTH_HIP(hipGetLastError());
  }
  return PTH_COMPLETE;
}

pth_status MgTree::Solve(MeshData &md, int *iters, std::vector<double> *history) {
  history->clear();
  double rms = 0.0;
  if (Rms(&rms) != PTH_COMPLETE) return PTH_FAIL;
  history->push_back(rms);
  int it = 0;
  do {
    if (VCycle() != PTH_COMPLETE) return PTH_FAIL;
    ++it;
    if (Rms(&rms) != PTH_COMPLETE) return PTH_FAIL;
    history->push_back(rms);
  } while (rms > P.residual_tolerance && it < P.max_iters);
  *iters = it;
  Mesh &m0 = *lv[0]->mesh;
  PTH_HIP(hipMemcpyAsync(md.store[fu].data, Data(0, kX), sizeof(double) * md.store[fu].ndata,
                         hipMemcpyDeviceToDevice, m0.stream));
  return PTH_COMPLETE;
}

MgTree *MgTreeCreate(Mesh *m, int fu, int frhs, int fD, const pth_mg_params &p, double D,
                     double alpha) {
  std::unique_ptr<MgTree> t(new MgTree());
  t->user = m;
  t->fu = fu;
  t->frhs = frhs;
  t->fD = fD;
  t->P = p;
  t->D = D;
  t->alpha = alpha;
  if (t->Build() != PTH_COMPLETE) return nullptr;
  return t.release();
}

void MgTreeDestroy(MgTree *t) { delete t; }

int MgTreeLevels(const MgTree *t) { return static_cast<int>(t->lv.size()); }

void MgTreeSetGraph(MgTree *t, bool on) { t->use_graph = on; }

pth_status MgTreeSetup(MgTree *t, MeshData &md) { return t->Setup(md); }

pth_status MgTreeSolve(MgTree *t, MeshData &md, int *iters, std::vector<double> *history) {
  return t->Solve(md, iters, history);
}

}  // namespace pth
