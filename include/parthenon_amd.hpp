// parthenon_amd.hpp — C++ mirror of the reference's task-level interface for the per-stage
// meshblock data path, implemented over the C ABI of parthenon_amd.h (header-only).
//
// A driver written against the reference (e.g. benchmarks/burgers/burgers_driver.cpp) calls
// the same names with the same argument meaning:
//   parthenon::SendBoundBufs<BT>, StartReceiveBoundBufs<BT>, ReceiveBoundBufs<BT>,
//   SetBounds<BT>, ProlongateBounds<BT>        src/bvals/comms/bvals_in_one.hpp:41-67
//   parthenon::Update::FluxDivergence, AverageIndependentData, UpdateIndependentData,
//   WeightedSumData                           src/interface/update.hpp:59-137
//   burgers_package::CalculateFluxes, CalculateDerived, EstimateTimestepMesh
//                                             benchmarks/burgers/burgers_package.hpp:24-30
//   MeshData<Real> from Mesh::mesh_data.GetOrAdd(name, partition)
//                                             src/interface/mesh_data.hpp:191, burgers_driver.cpp:76
// Error behaviour: TaskStatus::fail from the library becomes a std::runtime_error carrying the
// library's message (the reference's PARTHENON_REQUIRE_THROWS / PARTHENON_FAIL).
#ifndef PARTHENON_AMD_HPP_
#define PARTHENON_AMD_HPP_

#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <set>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "parthenon_amd.h"

namespace parthenon {

using Real = double;

enum class TaskStatus { complete, incomplete, iterate, fail };  // src/basic_types.hpp:59

enum class BoundaryType : int {  // src/basic_types.hpp:156
  any = PTH_BT_ANY,
  local = PTH_BT_LOCAL,
  nonlocal = PTH_BT_NONLOCAL,
  flxcor_send = PTH_BT_FLXCOR_SEND,
  flxcor_recv = PTH_BT_FLXCOR_RECV
};

struct Metadata {  // the cell-variable flags used on this path (src/interface/metadata.hpp)
  enum Flag : unsigned {
    FillGhost = PTH_FILL_GHOST,
    Independent = PTH_INDEPENDENT,
    WithFluxes = PTH_WITH_FLUXES,
    OneCopy = PTH_ONE_COPY,
    Vector = PTH_VECTOR
  };
};

// src/defs.hpp:127-137 BoundaryFace: the six faces of a block or of the domain
enum class BoundaryFace { inner_x1 = 0, outer_x1, inner_x2, outer_x2, inner_x3, outer_x3 };

namespace detail {
inline TaskStatus Check(pth_status s, const char *what) {
  if (s == PTH_FAIL) {
    const char *msg = pth_last_error();
    throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "unknown error"));
  }
  return static_cast<TaskStatus>(s);
}
}  // namespace detail

// ---- tasks (src/tasks/tasks.hpp) ----
// On one GPU a partition holds every block and the library orders device work on the mesh
// stream, so a task list runs its tasks in the order they were added (every dependency was
// added before the task that names it). A task returning incomplete is polled again, like
// Task::operator() re-enqueueing itself (tasks.hpp:93-118); fail and iterate end the list
// with an exception (iterative sublists are not offered: the solvers iterate inside the
// library).
struct TaskID {
  int id = -1;  // -1: no dependency
  TaskID() = default;
  explicit TaskID(int i) : id(i) {}
  TaskID operator|(const TaskID &o) const { return TaskID(id > o.id ? id : o.id); }
};

class TaskList {
 public:
  // AddTask(dependency, function, arguments...): the function is called with copies of the
  // arguments when the list runs
  template <class F, class... Args>
  TaskID AddTask(TaskID /*dependency*/, F &&f, Args &&...args) {
    tasks_.emplace_back(
        [f = std::forward<F>(f), a = std::make_tuple(std::forward<Args>(args)...)]() mutable {
          return std::apply(f, a);
        });
    return TaskID(static_cast<int>(tasks_.size()) - 1);
  }
  template <class F>
  TaskID AddTask(F &&f) {
    return AddTask(TaskID(), std::forward<F>(f));
  }
  int size() const { return static_cast<int>(tasks_.size()); }
  void Execute() {
    for (size_t q = 0; q < tasks_.size(); ++q) {
      TaskStatus s;
      int polls = 0;
      while ((s = tasks_[q]()) == TaskStatus::incomplete)
        if (++polls > kMaxPolls)
          throw std::runtime_error("TaskList: task " + std::to_string(q) + " never completed");
      if (s != TaskStatus::complete)
        throw std::runtime_error("TaskList: task " + std::to_string(q) +
                                 (s == TaskStatus::fail ? " failed" : " asked to iterate"));
    }
  }

 private:
  static constexpr int kMaxPolls = 1000;
  std::vector<std::function<TaskStatus()>> tasks_;
};

// TaskRegion / TaskCollection (tasks.hpp): regions run one after the other, the lists of a
// region (one per partition; one partition per GPU) one after the other
class TaskRegion {
 public:
  explicit TaskRegion(int n) : lists_(n) {}
  TaskList &operator[](int i) { return lists_.at(i); }
  int size() const { return static_cast<int>(lists_.size()); }
  void Execute() {
    for (auto &l : lists_) l.Execute();
  }

 private:
  std::vector<TaskList> lists_;
};

class TaskCollection {
 public:
  TaskRegion &AddRegion(int n) {
    regions_.emplace_back(new TaskRegion(n));
    return *regions_.back();
  }
  TaskStatus Execute() {
    for (auto &r : regions_) r->Execute();
    return TaskStatus::complete;
  }

 private:
  std::vector<std::unique_ptr<TaskRegion>> regions_;
};

// the reference's task-function wrapper; plain functions and lambdas are tasks here
#ifndef TF
#define TF(...) __VA_ARGS__
#endif

// AllReduce<T> (src/utils/reductions.hpp): on one GPU the local value is the global one
template <class T>
struct AllReduce {
  T val{};
  TaskStatus StartReduce(int = 0) { return TaskStatus::complete; }
  TaskStatus CheckReduce() { return TaskStatus::complete; }
};

class Mesh;

template <typename T>
class MeshData {
 public:
  MeshData(Mesh *mesh, pth_meshdata *h, std::string name)
      : mesh_(mesh), h_(h), name_(std::move(name)) {}
  Mesh *GetMeshPointer() const { return mesh_; }
  Mesh *GetParentPointer() const { return mesh_; }
  pth_meshdata *handle() const { return h_; }
  const std::string &StageName() const { return name_; }
  // host copies of a whole field (which: 0 data, 1 coarse buffer, 2 face fluxes)
  void Upload(const std::string &field, const std::vector<T> &host, int which = 0) {
    detail::Check(pth_field_upload(h_, field.c_str(), which, host.data(), host.size()), "Upload");
  }
  void Download(const std::string &field, std::vector<T> &host, int which = 0) const {
    detail::Check(pth_field_download(h_, field.c_str(), which, host.data(), host.size()),
                  "Download");
  }
  // MeshData::PackVariables / PackVariablesAndFluxes (mesh_data.hpp:349-430): the device view
  // of the named fields (a parthenon::SparsePack in HIP code, parthenon_amd_pack.hpp)
  pth_pack_view PackVariables(const std::vector<std::string> &vars) const {
    return Pack(vars, 0);
  }
  pth_pack_view PackVariablesAndFluxes(const std::vector<std::string> &vars) const {
    return Pack(vars, PTH_PACK_WITH_FLUXES);
  }
  pth_pack_view Pack(const std::vector<std::string> &vars, unsigned options) const {
    std::vector<const char *> names;
    for (const auto &v : vars) names.push_back(v.c_str());
    pth_pack *p = pth_pack_get(h_, names.data(), static_cast<int>(names.size()), options);
    if (!p) throw std::runtime_error(std::string("pack: ") + pth_last_error());
    pth_pack_view view;
    detail::Check(pth_pack_view_get(p, &view), "pack");
    return view;
  }

 private:
  Mesh *mesh_;
  pth_meshdata *h_;
  std::string name_;
};

// PDOpt (src/interface/sparse_pack_base.hpp:46)
enum class PDOpt { WithFluxes, Coarse, Flatten };

// MakePackDescriptor(md, vars, {}, options) (make_pack_descriptor.hpp:42-105) and
// Descriptor::GetPack (sparse_pack.hpp:71-108): a named-field descriptor whose packs are
// cached per MeshData; GetIdx(name) gives the field's PackIdx number.
class PackDescriptor {
 public:
  PackDescriptor(std::vector<std::string> vars, const std::set<PDOpt> &options)
      : vars_(std::move(vars)) {
    for (PDOpt o : options)
      opt_ |= o == PDOpt::WithFluxes ? PTH_PACK_WITH_FLUXES
              : o == PDOpt::Coarse   ? PTH_PACK_COARSE
                                     : PTH_PACK_FLATTEN;
  }
  pth_pack_view GetPack(const MeshData<Real> *md) const { return md->Pack(vars_, opt_); }
  int GetIdx(const std::string &name) const {
    for (size_t q = 0; q < vars_.size(); ++q)
      if (vars_[q] == name) return static_cast<int>(q);
    throw std::runtime_error("PackDescriptor: no variable " + name);
  }
  const std::vector<std::string> &vars() const { return vars_; }

 private:
  std::vector<std::string> vars_;
  unsigned opt_ = 0;
};

inline PackDescriptor MakePackDescriptor(const MeshData<Real> *, std::vector<std::string> vars,
                                         const std::set<PDOpt> &options = {}) {
  return PackDescriptor(std::move(vars), options);
}

// Mesh: a uniform or statically refined Cartesian mesh of equal meshblocks on one GPU.
class Mesh {
 public:
  struct Field {
    std::string name;
    int ncomp;
    unsigned flags;
  };
  Mesh(std::array<int, 3> nx, std::array<int, 3> block, int nghost,
       std::array<Real, 3> xmin = {-0.5, -0.5, -0.5}, std::array<Real, 3> xmax = {0.5, 0.5, 0.5},
       std::array<int, 3> periodic = {1, 1, 1}) {
    h_ = pth_mesh_create(nx.data(), block.data(), nghost, xmin.data(), xmax.data(),
                         periodic.data());
    if (!h_) throw std::runtime_error(std::string("Mesh: ") + pth_last_error());
  }
  ~Mesh() {
    data_.clear();
    if (h_) pth_mesh_destroy(h_);
  }
  Mesh(const Mesh &) = delete;
  Mesh &operator=(const Mesh &) = delete;

  // <parthenon/mesh> ix1_bc ... ox3_bc (src/mesh/mesh.cpp:1180-1185): "outflow",
  // "reflecting" or "fixed_face" (FixedFace with value 0, the poisson_gmg package's
  // condition) on a face of a non-periodic dimension (its default is outflow)
  void SetBoundaryCondition(BoundaryFace face, const std::string &name) {
    int type = name == "outflow"      ? PTH_BC_OUTFLOW
               : name == "reflecting" ? PTH_BC_REFLECT
               : name == "fixed_face" ? PTH_BC_FIXED_FACE
                                      : -1;
    if (type < 0) throw std::runtime_error("SetBoundaryCondition: unknown condition " + name);
    detail::Check(pth_mesh_set_boundary(h_, static_cast<int>(face), type),
                  "SetBoundaryCondition");
  }
  // <parthenon/static_refinementN> (src/mesh/mesh.cpp:1085-1116)
  void AddStaticRefinement(std::array<Real, 6> box, int level) {
    detail::Check(pth_mesh_refine_box(h_, box.data(), level), "AddStaticRefinement");
  }
  // static refinement of every block that intersects a sphere, as a union of per-block boxes
  void AddRefinementSphere(std::array<Real, 3> center, Real radius, int level) {
    detail::Check(pth_mesh_refine_sphere(h_, center.data(), radius, level),
                  "AddRefinementSphere");
  }
  // StateDescriptor::AddField for a cell-centred field
  void AddField(const std::string &name, int ncomp, unsigned flags) {
    detail::Check(pth_mesh_add_field(h_, name.c_str(), ncomp, flags), "AddField");
  }
  // Metadata::RegisterRefinementOps<ProlongateSharedLinear, RestrictAverage>() of a field
  // (metadata.hpp:555); ProlongateSharedMinMod is the default
  enum class ProlongationOp { SharedMinMod = PTH_PROLONG_MINMOD, SharedLinear = PTH_PROLONG_LINEAR };
  void RegisterRefinementOps(const std::string &field, ProlongationOp op) {
    detail::Check(pth_mesh_register_refinement_ops(h_, field.c_str(), static_cast<int>(op),
                                                   PTH_RESTRICT_AVERAGE),
                  "RegisterRefinementOps");
  }
  void Initialize() { detail::Check(pth_mesh_finalize(h_), "Initialize"); }

  // <parthenon/mesh> refinement = adaptive, numlevel, derefine_count (before Initialize)
  void SetAdaptive(int numlevel, int derefine_count = 10) {
    detail::Check(pth_mesh_set_adaptive(h_, numlevel, derefine_count), "SetAdaptive");
  }
  // <parthenon/refinementN>: method "derivative_order_1" or "derivative_order_2" on
  // component vector_i of field (AMRCriteria::MakeAMRCriteria, amr_criteria.cpp:70-80)
  void AddRefinementCriterion(const std::string &method, const std::string &field, int vector_i,
                              Real refine_tol = 0.5, Real derefine_tol = 0.05,
                              int max_level = 0) {
    const int m = method == "derivative_order_1" ? 1 : method == "derivative_order_2" ? 2 : -1;
    if (m < 0) throw std::invalid_argument("Invalid selection for refinement method: " + method);
    detail::Check(pth_mesh_add_refinement_criterion(h_, m, field.c_str(),
                                                    vector_i, refine_tol, derefine_tol,
                                                    max_level),
                  "AddRefinementCriterion");
  }
  // Mesh::LoadBalancingAndAdaptiveMeshRefinement (mesh-amr_loadbalance.cpp:324-351); the
  // result is Mesh::modified. MeshData objects from GetOrAdd stay valid.
  bool LoadBalancingAndAdaptiveMeshRefinement() {
    int modified = 0;
    detail::Check(pth_mesh_adapt(h_, &modified), "LoadBalancingAndAdaptiveMeshRefinement");
    return modified != 0;
  }
  // MeshRefinement::SetRefinement of block gid
  void SetRefinement(int gid, int tag) {
    detail::Check(pth_mesh_set_refinement(h_, gid, tag), "SetRefinement");
  }

  int nbtotal() const { return pth_mesh_nblocks(h_); }
  bool multilevel() const { return pth_mesh_multilevel(h_) != 0; }
  pth_mesh *handle() const { return h_; }

  // mesh_data.GetOrAdd(name, partition): one partition (every block) on one GPU
  std::shared_ptr<MeshData<Real>> &GetOrAdd(const std::string &name, int partition = 0) {
    if (partition != 0) throw std::runtime_error("GetOrAdd: one partition per GPU");
    auto it = data_.find(name);
    if (it == data_.end()) {
      pth_meshdata *h = pth_meshdata_get(h_, name.c_str());
      if (!h) throw std::runtime_error(std::string("GetOrAdd: ") + pth_last_error());
      it = data_.emplace(name, std::make_shared<MeshData<Real>>(this, h, name)).first;
    }
    return it->second;
  }
  int DefaultNumPartitions() const { return 1; }

 private:
  pth_mesh *h_ = nullptr;
  std::map<std::string, std::shared_ptr<MeshData<Real>>> data_;
};

// ---- refinement tagging (src/amr_criteria/refinement_package.hpp) ----
namespace Refinement {
template <typename T>
TaskStatus Tag(T *md);
template <>
inline TaskStatus Tag(MeshData<Real> *md) {
  return detail::Check(pth_refinement_tag(md->handle()), "Refinement::Tag");
}
}  // namespace Refinement

// ---- boundary communication tasks (src/bvals/comms/bvals_in_one.hpp:41-67) ----
template <BoundaryType BT>
TaskStatus SendBoundBufs(std::shared_ptr<MeshData<Real>> &md) {
  return detail::Check(pth_send_bound_bufs(md->handle(), static_cast<int>(BT)), "SendBoundBufs");
}
template <BoundaryType BT>
TaskStatus StartReceiveBoundBufs(std::shared_ptr<MeshData<Real>> &md) {
  return detail::Check(pth_start_receive_bound_bufs(md->handle(), static_cast<int>(BT)),
                       "StartReceiveBoundBufs");
}
template <BoundaryType BT>
TaskStatus ReceiveBoundBufs(std::shared_ptr<MeshData<Real>> &md) {
  return detail::Check(pth_receive_bound_bufs(md->handle(), static_cast<int>(BT)),
                       "ReceiveBoundBufs");
}
template <BoundaryType BT>
TaskStatus SetBounds(std::shared_ptr<MeshData<Real>> &md) {
  return detail::Check(pth_set_bounds(md->handle(), static_cast<int>(BT)), "SetBounds");
}
template <BoundaryType BT>
TaskStatus ProlongateBounds(std::shared_ptr<MeshData<Real>> &md) {
  return detail::Check(pth_prolongate_bounds(md->handle(), static_cast<int>(BT)),
                       "ProlongateBounds");
}
// aliases of bvals_in_one.hpp:50-61
inline TaskStatus SendBoundaryBuffers(std::shared_ptr<MeshData<Real>> &md) {
  return SendBoundBufs<BoundaryType::any>(md);
}
inline TaskStatus ReceiveBoundaryBuffers(std::shared_ptr<MeshData<Real>> &md) {
  return ReceiveBoundBufs<BoundaryType::any>(md);
}
inline TaskStatus SetBoundaries(std::shared_ptr<MeshData<Real>> &md) {
  return SetBounds<BoundaryType::any>(md);
}
inline TaskStatus ProlongateBoundaries(std::shared_ptr<MeshData<Real>> &md) {
  return ProlongateBounds<BoundaryType::any>(md);
}
// physical boundary conditions (src/bvals/boundary_conditions.hpp:40-50)
inline TaskStatus ApplyBoundaryConditionsOnCoarseOrFineMD(std::shared_ptr<MeshData<Real>> &md,
                                                         bool coarse) {
  return detail::Check(pth_apply_boundary_conditions(md->handle(), coarse ? 1 : 0),
                       "ApplyBoundaryConditionsOnCoarseOrFineMD");
}
inline TaskStatus ApplyBoundaryConditionsMD(std::shared_ptr<MeshData<Real>> &md) {
  return ApplyBoundaryConditionsOnCoarseOrFineMD(md, false);
}
// flux-correction tasks of bvals_in_one.hpp:69-80
inline TaskStatus StartReceiveFluxCorrections(std::shared_ptr<MeshData<Real>> &md) {
  return StartReceiveBoundBufs<BoundaryType::flxcor_recv>(md);
}
inline TaskStatus LoadAndSendFluxCorrections(std::shared_ptr<MeshData<Real>> &md) {
  return SendBoundBufs<BoundaryType::flxcor_send>(md);
}
inline TaskStatus ReceiveFluxCorrections(std::shared_ptr<MeshData<Real>> &md) {
  return ReceiveBoundBufs<BoundaryType::flxcor_recv>(md);
}
inline TaskStatus SetFluxCorrections(std::shared_ptr<MeshData<Real>> &md) {
  return SetBounds<BoundaryType::flxcor_recv>(md);
}

// ---- Update:: tasks (src/interface/update.hpp) ----
// AddBoundaryExchangeTasks<any> (src/bvals/comms/boundary_communication.cpp:405-445): send,
// receive, set; on a multilevel mesh the coarse-buffer conditions and the prolongation; then
// the physical boundary conditions
template <BoundaryType BT>
TaskID AddBoundaryExchangeTasks(TaskID dependency, TaskList &tl,
                                std::shared_ptr<MeshData<Real>> &md, bool multilevel) {
  static_assert(BT == BoundaryType::any, "AddBoundaryExchangeTasks: BoundaryType::any only");
  auto send = tl.AddTask(dependency, SendBoundBufs<BT>, md);
  auto recv = tl.AddTask(dependency, ReceiveBoundBufs<BT>, md);
  auto set = tl.AddTask(send | recv, SetBounds<BT>, md);
  auto pro = set;
  if (multilevel) {
    auto cbound = tl.AddTask(set, ApplyBoundaryConditionsOnCoarseOrFineMD, md, true);
    pro = tl.AddTask(cbound, ProlongateBounds<BT>, md);
  }
  return tl.AddTask(pro, ApplyBoundaryConditionsOnCoarseOrFineMD, md, false);
}

namespace Update {
template <typename T>
TaskStatus FluxDivergence(T *in, T *dudt) {
  return detail::Check(pth_flux_divergence(in->handle(), dudt->handle()), "FluxDivergence");
}
template <typename T>
TaskStatus WeightedSumDataIndependent(T *in1, T *in2, const Real w1, const Real w2, T *out) {
  return detail::Check(
      pth_weighted_sum_independent(in1->handle(), in2->handle(), w1, w2, out->handle()),
      "WeightedSumData");
}
template <typename T>
TaskStatus UpdateIndependentData(T *in, T *dudt, const Real dt, T *out) {
  return WeightedSumDataIndependent(in, dudt, 1.0, dt, out);
}
template <typename T>
TaskStatus AverageIndependentData(T *c1, T *c2, const Real wgt1) {
  return WeightedSumDataIndependent(c1, c2, wgt1, (1.0 - wgt1), c1);
}
}  // namespace Update

// ---- multigrid Poisson solver (src/solvers/mg_solver.hpp:37-521) ----
namespace solvers {

// Field types name fields the way the reference's VARIABLE(ns, varname) types do
// (src/interface/make_pack_descriptor.hpp): a type with a static name().
#define PARTHENON_AMD_VARIABLE(ns, varname)                     \
  struct varname {                                              \
    static std::string name() { return #ns "." #varname; }     \
  }

// solver_utils (src/solvers/solver_utils.hpp:151-301) over every block of md
namespace utils {
template <class in_t, class out_t>
TaskStatus CopyData(std::shared_ptr<MeshData<Real>> &md) {  // :151-178, ghosts included
  return detail::Check(
      pth_solver_copy_data(md->handle(), in_t::name().c_str(), out_t::name().c_str()),
      "CopyData");
}
template <class a_t, class b_t, class out_t>
TaskStatus AddFieldsAndStore(std::shared_ptr<MeshData<Real>> &md, Real wa = 1.0,
                             Real wb = 1.0) {  // :180-222, out = wa a + wb b
  return detail::Check(pth_solver_add_fields_and_store(md->handle(), a_t::name().c_str(),
                                                       b_t::name().c_str(),
                                                       out_t::name().c_str(), wa, wb),
                       "AddFieldsAndStore");
}
template <class var_t>
TaskStatus SetToZero(std::shared_ptr<MeshData<Real>> &md) {  // :224-250, interior cells
  return detail::Check(pth_solver_set_to_zero(md->handle(), var_t::name().c_str()),
                       "SetToZero");
}
template <class a_t, class b_t>
TaskStatus DotProductLocal(std::shared_ptr<MeshData<Real>> &md, AllReduce<Real> *adotb) {
  Real v = 0;
  TaskStatus s = detail::Check(
      pth_solver_dot_product(md->handle(), a_t::name().c_str(), b_t::name().c_str(), &v),
      "DotProduct");
  adotb->val += v;
  return s;
}
// DotProduct (:279-301): zero the result, add this partition's sum (one partition per GPU)
template <class a_t, class b_t>
TaskID DotProduct(TaskID dependency_in, TaskList &tl, AllReduce<Real> *adotb,
                  std::shared_ptr<MeshData<Real>> &md) {
  auto zero = tl.AddTask(dependency_in, [adotb]() {
    adotb->val = 0.0;
    return TaskStatus::complete;
  });
  return tl.AddTask(zero, DotProductLocal<a_t, b_t>, md, adotb);
}
}  // namespace utils

// The poisson_gmg example's PoissonEquation (example/poisson_gmg/poisson_equation.hpp:41-205)
// with a constant D and the package's diagonal_alpha. do_flux_cor: flux correction at
// fine-coarse faces; the operator here runs on single-level meshes, where it changes nothing.
class PoissonEquation {
 public:
  Real D = 1.0;               // constant D, or
  std::string D_field;        // the face field D: a 3-component field, component d = D on
                              // the lower face along d (the reference's Metadata::Face D)
  Real diagonal_alpha = 0.0;
  bool do_flux_cor = false;
  pth_poisson_equation Coefficients() const {
    return {D, diagonal_alpha, D_field.empty() ? nullptr : D_field.c_str()};
  }
  // Ax<x_t, out_t> (:41-53): out = A x on every block's interior; x's ghosts must be filled
  template <class x_t, class out_t>
  TaskID Ax(TaskList &tl, TaskID depends_on, std::shared_ptr<MeshData<Real>> &md) const {
    // the task keeps its own copy (the face field's name outlives this object)
    return tl.AddTask(depends_on, [eq = *this, md]() {
      const pth_poisson_equation e = eq.Coefficients();
      return detail::Check(
          pth_poisson_ax(md->handle(), &e, x_t::name().c_str(), out_t::name().c_str()), "Ax");
    });
  }
};

// MGParams (mg_solver.hpp:37-58): max_iters and residual_tolerance (RMS residual
// sqrt(sum r^2 / ncells), the reference's stop test, :147-158) as the reference; the smoother
// is weighted Jacobi with an explicit weight and sweep counts ("jacobi") or the reference's
// "SRJ1" / "SRJ2" / "SRJ3" stage weights (:312-345); mixed_precision (FP32 V-cycles inside
// FP64 iterative refinement) applies to the uniform N^3 cube with fixed_face faces only.
struct MGParams {
  int max_iters = 1000;
  Real residual_tolerance = 1.e-12;
  Real omega = 2.0 / 3.0;
  int pre_sweeps = 2, post_sweeps = 2;
  bool mixed_precision = true;  // FP32 V-cycles inside FP64 iterative refinement
  int inner_cycles = 2;          // FP32 V-cycles per FP64 refinement step
  std::string smoother = "jacobi";
};

// MGSolver<u, rhs, equations> (mg_solver.hpp:73-520) over the blocks of a MeshData: u and rhs
// are field types (static name()), equations carries the operator's coefficients
// (Coefficients(), as PoissonEquation). Faces are fixed_face (the poisson_gmg package's
// FixedFace) or periodic; the mesh may be statically refined, 1-D, 2-D or 3-D (a uniform
// 3-D N^3 cube with fixed_face faces runs on whole-level arrays, any other mesh on a
// hierarchy of block meshes in FP64).
template <class u, class rhs, class equations>
class MGSolver {
 public:
  explicit MGSolver(MGParams params_in = MGParams(), equations eq_in = equations(),
                    std::string container = "base")
      : params_(params_in), eqs_(eq_in), container_(std::move(container)) {}
  ~MGSolver() {
    if (h_) pth_mg_destroy(h_);
  }
  MGSolver(const MGSolver &) = delete;
  MGSolver &operator=(const MGSolver &) = delete;

  // AddSetupTasks (:188-203): gathers rhs and the current u of the partition's blocks
  template <class TL_t>
  TaskID AddSetupTasks(TL_t &tl, TaskID dependence, int partition, Mesh *pmesh) {
    auto &md = pmesh->GetOrAdd(container_, partition);
    return tl.AddTask(dependence, [this, pmesh, md]() {
      if (!h_) Create(pmesh);
      return detail::Check(pth_mg_setup(h_, md->handle()), "MGSolver::AddSetupTasks");
    });
  }
  // AddTasks (:119-161): V-cycles until the RMS residual is below residual_tolerance (at
  // least one, at most max_iters); u's interior and ghosts hold the solution afterwards
  TaskID AddTasks(TaskList &tl, TaskID dependence, Mesh *pmesh, int partition) {
    auto &md = pmesh->GetOrAdd(container_, partition);
    return tl.AddTask(dependence, [this, md]() {
      if (!h_) throw std::runtime_error("MGSolver::AddTasks: AddSetupTasks has not run");
      TaskStatus s = detail::Check(pth_mg_solve(h_, md->handle()), "MGSolver::AddTasks");
      detail::Check(pth_mg_result(h_, &final_iteration_, &final_residual_), "MGSolver");
      return s;
    });
  }
  Real GetFinalResidual() const { return final_residual_; }
  int GetFinalIterations() const { return final_iteration_; }
  // RMS residual before every V-cycle of the last solve and after its last
  std::vector<Real> GetResidualHistory() const {
    std::vector<Real> h(1024);
    const int n = h_ ? pth_mg_history(h_, h.data(), static_cast<int>(h.size())) : 0;
    h.resize(n < 1024 ? n : 1024);
    return h;
  }

 private:
  void Create(Mesh *pmesh) {
    pth_mg_params p{};
    p.max_iters = params_.max_iters;
    p.residual_tolerance = params_.residual_tolerance;
    p.precision = params_.mixed_precision ? 0 : 1;
    p.inner_cycles = params_.inner_cycles;
    p.omega = params_.omega;
    p.pre_sweeps = params_.pre_sweeps;
    p.post_sweeps = params_.post_sweeps;
    p.smoother = params_.smoother == "SRJ1"   ? 1
                 : params_.smoother == "SRJ2" ? 2
                 : params_.smoother == "SRJ3" ? 3
                 : params_.smoother == "jacobi"
                     ? 0
                     : throw std::runtime_error("MGSolver: unknown smoother " + params_.smoother);
    const pth_poisson_equation e = eqs_.Coefficients();
    h_ = pth_mg_create(pmesh->handle(), u::name().c_str(), rhs::name().c_str(), &p, &e);
    if (!h_) throw std::runtime_error(std::string("MGSolver: ") + pth_last_error());
  }
  MGParams params_;
  equations eqs_;
  std::string container_;
  pth_mg *h_ = nullptr;
  Real final_residual_ = 0;
  int final_iteration_ = 0;
};

// The standalone solver on the uniform n^3 grid of example/poisson_gmg ([0,1]^3, FixedFace
// 0) without a mesh: set the right-hand side, Solve(), read the solution; host arrays in
// [k][j][i]; residual_tolerance is relative here, ||b - A u|| / ||b||.
struct UniformMGParams {
  int max_iters = 1000;
  Real residual_tolerance = 1.e-10;
  Real omega = 2.0 / 3.0;
  int pre_sweeps = 2, post_sweeps = 2;
  bool mixed_precision = true;
  int inner_cycles = 2;
  bool column_sweeps = true;     // Jacobi by z columns where the level size allows
};

class UniformMGSolver {
 public:
  UniformMGSolver(int n, UniformMGParams p = UniformMGParams()) : n_(n), p_(p) {
    h_ = pth_gmg_create(n, p.mixed_precision ? 0 : 1, p.omega, p.pre_sweeps, p.post_sweeps);
    if (!h_) throw std::runtime_error(std::string("UniformMGSolver: ") + pth_last_error());
    detail::Check(pth_gmg_set_inner_cycles(h_, p.mixed_precision ? p.inner_cycles : 1),
                  "UniformMGSolver");
    detail::Check(pth_gmg_set_column_sweeps(h_, p.column_sweeps ? 1 : 0), "UniformMGSolver");
  }
  ~UniformMGSolver() {
    if (h_) pth_gmg_destroy(h_);
  }
  UniformMGSolver(const UniformMGSolver &) = delete;
  UniformMGSolver &operator=(const UniformMGSolver &) = delete;
  void SetRhs(const std::vector<Real> &b) {
    detail::Check(pth_gmg_set_rhs(h_, b.data(), b.size()), "SetRhs");
  }
  void SetRhsBlobs(std::uint64_t seed) {
    detail::Check(pth_gmg_set_rhs_blobs(h_, seed), "SetRhsBlobs");
  }
  int Solve() {
    int it = 0;
    detail::Check(pth_gmg_solve(h_, p_.residual_tolerance, p_.max_iters, &it, &final_residual_),
                  "Solve");
    final_iteration_ = it;
    return it;
  }
  void GetSolution(std::vector<Real> &u) const {
    u.resize(static_cast<size_t>(n_) * n_ * n_);
    detail::Check(pth_gmg_get_solution(h_, u.data(), u.size()), "GetSolution");
  }
  Real GetFinalResidual() const { return final_residual_; }
  int GetFinalIterations() const { return final_iteration_; }

 private:
  int n_;
  UniformMGParams p_;
  pth_gmg *h_ = nullptr;
  Real final_residual_ = 0;
  int final_iteration_ = 0;
};
}  // namespace solvers

}  // namespace parthenon

// ---- the Burgers benchmark package (benchmarks/burgers/burgers_package.hpp) ----
namespace burgers_package {
using parthenon::MeshData;
using parthenon::Real;
using parthenon::TaskStatus;
enum class ReconType { WENO5 = 0, Linear = 1 };

inline TaskStatus CalculateFluxes(MeshData<Real> *md, ReconType recon = ReconType::WENO5) {
  return parthenon::detail::Check(
      pth_burgers_calculate_fluxes(md->handle(), static_cast<int>(recon)), "CalculateFluxes");
}
inline void CalculateDerived(MeshData<Real> *md) {
  parthenon::detail::Check(pth_burgers_calculate_derived(md->handle()), "CalculateDerived");
}
inline Real EstimateTimestepMesh(MeshData<Real> *md, Real cfl) {
  Real dt = 0;
  parthenon::detail::Check(pth_burgers_estimate_timestep(md->handle(), cfl, &dt),
                           "EstimateTimestepMesh");
  return dt;
}
}  // namespace burgers_package

#endif  // PARTHENON_AMD_HPP_
