// Internal host-side structures of the parthenon_amd library.
//
// Layout in HBM (one GPU, all blocks in one MeshData):
//   field data   [block][comp][k][j][i]         i fastest, (n+2g)^3 brick per component
//   coarse data  [block][comp][ck][cj][ci]      reference coarse_s, (n/2+2g)^3 per component
//   face fluxes  [dir][block][comp][k][j][i]    x-face i stored at index i
//   comm buffers one flat arena per field name, each channel a contiguous slice laid out
//                [comp][k][j][i] exactly like the reference buffer order
//   region table one 96-byte record per (block, neighbour, field) boundary region, walked
//                by a single persistent kernel per pack / unpack / restrict / prolong call
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/parthenon_amd.h"

namespace pth {

void SetError(const std::string &msg);
const char *ErrorText();

#define PTH_HIP(call)                                                              \
  do {                                                                             \
    hipError_t err_ = (call);                                                      \
    if (err_ != hipSuccess) {                                                      \
      ::pth::SetError(std::string(#call) + ": " + hipGetErrorString(err_));        \
      return PTH_FAIL;                                                             \
    }                                                                              \
  } while (0)

// ---------------------------------------------------------------- topology ---------------
struct Lattice {  // a block position: level and integer coordinates on that level's lattice
  int level = 0;
  std::int64_t x[3] = {0, 0, 0};
  bool operator<(const Lattice &o) const {
    return std::tie(level, x[2], x[1], x[0]) < std::tie(o.level, o.x[2], o.x[1], o.x[0]);
  }
  bool operator==(const Lattice &o) const {
    return level == o.level && x[0] == o.x[0] && x[1] == o.x[1] && x[2] == o.x[2];
  }
};

struct Link {             // one neighbour of a block (reference NeighborBlock, CC subset)
  int gid;                // neighbour block id
  int level;              // neighbour level
  std::int64_t seen[3];   // neighbour coordinates as seen from this block (not wrapped)
  int off[3];             // offset direction, each in {-1, 0, 1}
  int loc_index() const { return (off[0] + 1) + 3 * (off[1] + 1) + 9 * (off[2] + 1); }
  int reverse_index() const { return (1 - off[0]) + 3 * (1 - off[1]) + 9 * (1 - off[2]); }
  bool is_cell() const { return off[0] == 0 && off[1] == 0 && off[2] == 0; }
};

struct Block {
  Lattice at;
  double xmin[3], xmax[3];
  double dx[3];      // fine cell widths
  double x0[3];      // fine coordinate origin: xmin - g dx  (cell centre = x0 + (i+.5) dx)
  double cdx[3];     // coarse cell widths
  double cx0[3];     // coarse coordinate origin
  double vol;        // fine cell volume
  double area[3];    // fine face areas
  bool coarser_neighbor = false;
  std::vector<Link> links;
};

struct Extent {  // index ranges of a block array per dimension (reference IndexShape)
  int n[3], g[3];
  int ent(int d) const { return n[d] + 2 * g[d]; }
  int lo(int d) const { return g[d]; }
  int hi(int d) const { return g[d] + n[d] - 1; }
  std::int64_t cells() const { return (std::int64_t)ent(0) * ent(1) * ent(2); }
};

struct Box {  // inclusive index box [lo, hi] per dimension
  int lo[3], hi[3];
  std::int64_t count() const {
    return (std::int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  }
};

enum class Side { kSendBoundary, kRecvBoundary, kSendInterior, kRecvInterior };

class Topology {
 public:
  int nx[3], bnx[3], ng;
  double xmin[3], xmax[3];
  bool periodic[3], active[3];
  int bc[6];  // pth_bc of each domain face (periodic dimensions: PTH_BC_PERIODIC)
  int ndim = 3;
  std::int64_t nbase[3];
  bool multilevel = false;
  std::set<Lattice> leaf, inner;
  std::vector<Block> blocks;
  std::map<Lattice, int> id;
  Extent fine, coarse;

  std::string Init(const int *nx_, const int *bnx_, int g, const double *lo, const double *hi,
                   const int *per);
  void RefineBox(const double box[6], int level);
  void RefineSphere(const double c[3], double r, int level);
  std::string Finalize();
  // Adaptive refinement of a finalized tree (leaf / inner sets; Finalize rebuilds the blocks).
  // Split one leaf with 2:1 nesting (reference Tree::Refine); returns the blocks added.
  int Refine(const Lattice &a);
  // Merge the children of `parent` back into it if they are all leaves and no same-level
  // neighbour of a child is an inner node (reference Tree::Derefine); returns blocks removed.
  int Derefine(const Lattice &parent);
  // log2 of the largest power of two dividing every active dimension's base block count (0 if
  // a dimension is inactive): the reference's root level (Forest::HyperRectangular).
  int RootLevel() const;
  // true when every neighbour link joins blocks at most one level apart
  bool Nested() const;
  // Index box of one boundary region (reference CalcIndices, cell-centred fields).
  Box Indices(const Block &b, const Link &l, Side side, bool prores) const;

 private:
  std::int64_t Span(int d, int level) const { return nbase[d] << level; }
  bool Fold(Lattice &a) const;
  std::vector<Lattice> Children(const Lattice &a) const;
  int Split(const Lattice &a);
  void Ensure(const Lattice &a);
  double Position(int d, std::int64_t index, int edge, int level) const;
  void Neighbors(int gid);
};

// ---------------------------------------------------------------- device tables ----------
// One boundary region processed by the persistent kernels. For pack/unpack, `a` is the field
// slab (data or coarse) of the block and `b` the start of its buffer slice. For restriction
// and prolongation, `a` is the fine slab and `b` the coarse slab of the same block.
struct Region {
  double *a;
  double *b;
  std::int64_t start;     // first element in the flattened element space of the table
  int lo[3];              // index box origin (i, j, k) in the index space of the box
  int n[3];               // box extents (i, j, k)
  int nc;                 // components
  int gid;                // block id (coordinates for prolongation)
  int astride[3];         // element strides of `a` for (j, k, comp)
  int bstride[3];         // element strides of `b` for (j, k, comp) (restrict/prolong)
  std::uint64_t magic[3]; // ceil(2^40 / n[d]): q = (x * magic) >> 40 for x < 2^30 / n[d]
};
static_assert(sizeof(Region) % 8 == 0, "region record alignment");

struct DeviceBlockGeom {  // per-block coordinates for prolongation
  double x0[3], dx[3], cx0[3], cdx[3];
};

struct BlockCoefs {  // per-block uniform-Cartesian metrics used by the stencils
  double dx[3];
  double area[3];
  double vol;
  double pad;
};

// Same-level ghost-shell exchange (ShellKernel, pth_bvals.hip): per (field, component), the
// brick of block 0 and the element stride between consecutive blocks' bricks.
struct ShellBrick {
  double *base;
  std::int64_t stride;
};
constexpr int kShellBricks = 32;  // bricks per launch (kernel arguments)
struct ShellArgs {
  double *brick[kShellBricks];
  std::int64_t stride[kShellBricks];
  const int *same;   // [block][27] same-level neighbour or -1
  const int *order;  // blocks in Morton order of their positions
  int nblocks;
};

struct Item {  // a run of consecutive elements over up to kItemRegions consecutive regions
  int region;    // first region
  int nreg;      // regions touched
  int begin;     // element offset within the first region
  int count;     // elements
};
constexpr int kItemRegions = 8;
constexpr int kItemElems = 4096;
// restriction / prolongation items: one thread does a few coarse cells (each with 8 dependent
// fine loads or stores), so these tables are cut finer to keep enough workgroups in flight
constexpr int kProResItemElems = 512;

struct Table {
  std::vector<Region> host;
  std::vector<Item> items;
  Region *dev = nullptr;
  Item *dev_items = nullptr;
  std::int64_t total = 0;  // total elements
  int item_elems = kItemElems;  // elements per work item (at most kItemElems)
  void Free();
  pth_status Upload();
};

// Flux correction (reference flxcor_send / flxcor_recv, boundary_communication.cpp:454-461; the
// face-element index boxes of bnd_info.cpp:69-101, 106-212, 273-305 and 387-403; the neighbour
// selection of loop_utils.hpp:133-162): one patch per (fine block, coarser face neighbour).
// Its coarse face cells are the 2x2 (2-D: 2, 1-D: 1) area-weighted averages of the fine
// block's face fluxes; they replace the coarse neighbour's fluxes on the matching part of its
// face. Values live in a correction store holding one image per corrected block face,
// laid out [comp][t2][t1] over that face's interior cells (t1 < t2 the transverse dimensions).
struct FluxPatch {
  int fine, coarse, d;
  int fface;          // fine face index along d (g or g + n)
  int cface;          // receiver face index along d
  int ft[2];          // fine transverse start (t1, t2)
  int ct[2];          // receiver transverse start (t1, t2)
  int n[2];           // coarse face cells along t1, t2
  int step[2];        // fine cells per coarse cell along t1, t2 (2 if active, else 1)
  int rows, comp;     // store strides between t2 rows and between components
  std::int64_t dst;   // store offset of coarse face cell (0, 0), component 0
  double area;        // fine face area (reference Coordinates::Volume<F(d+1)>)
  double pad;
};
static_assert(sizeof(FluxPatch) % 8 == 0, "patch record alignment");

FluxPatch PatchGeometry(const Topology &topo, int fine, int coarse, int d, int sgn);

struct FluxCorrection {  // per field with fluxes
  bool built = false;
  int ncomp = 0;
  std::vector<FluxPatch> host;
  std::vector<int> face;   // [block][6] store offset of face (d, side) image, or -1
  FluxPatch *dev = nullptr;
  int *dev_face = nullptr;
  double *store = nullptr;
  std::int64_t store_n = 0;
  int max_cells = 0;       // largest patch (coarse face cells)
  void Free();
};

struct FieldSpec {
  std::string name;
  int ncomp;
  unsigned flags;
  int prolong = PTH_PROLONG_MINMOD;  // registered shared-prolongation operator
};

struct FieldStore {
  double *data = nullptr, *coarse = nullptr, *flux = nullptr;
  std::int64_t ndata = 0, ncoarse = 0, nflux = 0;
  bool owned = true;
};

class Mesh;

// Adaptive refinement settings and per-block state (reference MeshRefinement members and the
// <parthenon/refinementN> criteria).
struct AmrCriterion {
  int method;     // 1 derivative_order_1, 2 derivative_order_2
  int field;      // field index
  int comp;
  double refine_tol, derefine_tol;
  int max_level;  // compared with the logical level root_level + level
};

struct Adaptive {
  bool on = false;
  int numlevel = 1, deref_threshold = 10;
  std::vector<AmrCriterion> crit;
  std::vector<int> flag, count;  // refine_flag_, deref_count_ per block
  double *maxd = nullptr;        // device: [block][criterion] derivative maxima
  std::int64_t maxd_n = 0;
};

// One box of ghost cells of one block and field that the physical boundary conditions fill
// (reference GenericBC, boundary_conditions_generic.hpp:174-255). Along each dimension every
// cell of the box lies in the same class (inner ghosts, interior, outer ghosts), so each cell
// has one source cell src[d] = a[d] + b[d] * idx[d]: b = 1, a = 0 off the condition's faces;
// b = 0, a = first / last interior index for outflow; b = -1, a = 2 * ref -/+ 1 for
// reflection. Sources lie in no condition's ghost region, so one launch equals the
// reference's face-by-face passes; component c < 3 of a vector field is negated when
// reflected along c (bit c of flip).
struct BcBox {
  double *slab;  // component 0 of the block's data or coarse brick
  int lo[3], n[3];
  int a[3], b[3];
  int e[3];      // brick extents (i, j, k)
  int ncomp;
  int flip;      // reflecting dimensions
  int vec;       // 1 for a Metadata::Vector field
  int fixed;     // FixedFace-0 dimensions (every component: 2 * 0 - value, in face order)
  int pad;
};
static_assert(sizeof(BcBox) % 8 == 0, "bc box alignment");

struct BcTable {
  std::vector<BcBox> host;
  BcBox *dev = nullptr;
  void Free();
  pth_status Upload();
};

struct BoundCache {  // per (MeshData, bound type): the pack / unpack / pro-res tables
  bool built = false;
  Table pack, unpack, restrict_send, restrict_set, prolong;
  Table prolong_linear;  // prolongation regions of fields registered with the linear operator
  BcTable bc[2];  // physical boundary conditions on data (0) and coarse buffers (1)
  std::int64_t nbuffers = 0;
  // direct-exchange index pairs without the same-level channels of field rest_field (filled
  // by the block sweep itself); built on first use
  Table pack_rest, unpack_rest;
  int rest_field = -1;
  // direct-exchange pairs of the channels between blocks of different levels, with the
  // same-level channels of every field served by the ghost-shell kernel; built on first use
  Table pack_cross, unpack_cross;
  std::vector<ShellBrick> shell;
  int *shell_order = nullptr;  // device: block ids in Morton order
  bool cross_built = false;
};

// <parthenon/time> limits of EvolutionDriver with its defaults (src/driver/driver.hpp:66-94)
struct ClockLimits {
  double tlim = HUGE_VAL;
  double dt_init = DBL_MAX;
  double dt_user = DBL_MAX;
  double dt_force = -DBL_MAX;
  double dt_floor = DBL_MIN;
  double dt_ceil = DBL_MAX;
  int dt_init_force = 0;
};

// One pack over a MeshData (reference SparsePack / VariablePack / VariableFluxPack): a device
// table of (block, variable) brick pointers, the face-flux bricks with PTH_PACK_WITH_FLUXES,
// and the first / last pack index of every requested field. Built on first request and cached
// by (fields, options) like SparsePackCache (sparse_pack_base.cpp:319-346); a remesh (which
// moves every brick) makes the next request rebuild it.
struct PackEntry {
  void *dev = nullptr;              // one allocation: data pointers, flux pointers, bounds
  pth_pack_view view{};
  std::vector<int> bounds;          // host copy, [2][nfields]
  int generation = -1;              // mesh generation the pointers belong to
  void Free();
};

struct MeshData {
  Mesh *mesh;
  std::string name;
  std::vector<FieldStore> store;  // indexed like Mesh::fields
  BoundCache cache[10];
  std::map<std::string, PackEntry> packs;  // keyed by options and field names
  ~MeshData();
};

// channel key: sender gid, receiver gid, field index, location index
using Channel = std::tuple<int, int, int, int>;

enum class BufState { kStale, kSending, kReceived };

class Mesh {
 public:
  Topology topo;
  std::vector<FieldSpec> fields;
  std::map<std::string, std::unique_ptr<MeshData>> data;
  std::vector<FieldStore> one_copy;             // shared storage of OneCopy fields
  std::vector<double *> arena;                  // comm buffer arena seen by each field
  std::vector<std::int64_t> arena_size;
  double *arena_base = nullptr;                 // the single arena of every FillGhost field
  std::int64_t arena_total = 0;
  std::map<Channel, std::int64_t> channel_offset;  // element offset within the field arena
  BufState state[10] = {};
  DeviceBlockGeom *geom = nullptr;
  BlockCoefs *coef = nullptr;
  double *face = nullptr;                       // Burgers face-state scratch
  int face_V = 0;
  // device-resident Burgers driver state: clock (kClockSlots doubles), its limits and the
  // captured cycle
  double *dtvar = nullptr;
  ClockLimits run_lim;
  hipStream_t run_stream = nullptr;
  hipEvent_t order_ev[2] = {nullptr, nullptr};  // mesh stream <-> driver stream ordering
  hipStream_t side_stream = nullptr;            // driver: the exchange beside the stencils
  hipEvent_t fork_ev[2] = {nullptr, nullptr};
  int run_overlap = 1;                          // driver: overlap exchange and stencils
  int rf_ppw[3] = {1, 1, 1};                    // driver: planes per workgroup, recon parts
  int run_sweep = 1;                            // driver: one block-sweep launch per stage
  bool form_auto = true;                        // driver: pick sweep / overlap / serial by timing
  bool form_exact = false;                      // driver: timing picks among bit-exact forms only
  bool form_tuned = false;
  int sweep_waves = 4;                          // driver: sweep register budget (waves/SIMD)
  // driver, sweep form on multilevel 16^3 meshes: each stage's sweep split into the blocks
  // without channels to other levels or physical faces (list A) and the others (list B), the
  // previous stage's cross-level exchange running beside sweep A
  int sweep_split = 1;
  int *split_list = nullptr;  // device: list A, then list B
  int split_nA = 0, split_nB = 0, split_gen = -1;
  int sweep_spread = 1;                         // driver: resident sweep workgroups / spread
  bool rf_tuned = false;
  bool rf_auto = true;                          // driver: tuning chosen by timing (not set)
  int tuned_blocks = 0;                         // driver: block count the choices were timed on
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  double run_cfl = 0.0;
  int run_recon = 0;
  bool run_ready = false;
  hipStream_t stream = nullptr;
  int pack_mode = PTH_PACK_PERSISTENT;
  int generation = 0;      // counts remeshes (packs built before one are rebuilt)
  bool force_coarse = false;  // coarse buffers on a single-level mesh (multigrid levels)
  bool planned = false;    // host topology built
  bool finalized = false;  // device storage, tables and buffers built
  double *scratch = nullptr;                    // reduction partials
  double *pinned = nullptr;                     // host-pinned readback slot
  std::int64_t scratch_n = 0;
  std::map<int, FluxCorrection> flxcor;         // by field index (fields with fluxes)
  BufState flxcor_state = BufState::kStale;

  ~Mesh();
  int FieldIndex(const std::string &n) const;
  pth_status Plan();
  pth_status Finalize();
  MeshData *Get(const std::string &name);
  pth_status Allocate(MeshData &md);
  pth_status BuildCache(MeshData &md, int bt);
  pth_status BuildBoundaryConditions(MeshData &md, BoundCache &c);
  // pack_rest / unpack_rest of md's exchange cache for field f; false if the exchange has no
  // gather form (then the caller keeps the full exchange)
  pth_status BuildRestExchange(MeshData &md, int f, bool *ok);
  // same-level channels of every field by ghost shells, the others in pack_cross /
  // unpack_cross; false in ok where the shell kernel does not apply
  pth_status BuildCrossExchange(MeshData &md, bool *ok);
  // per block, the same-level neighbour in each of the 27 directions (or -1), on the device
  int *same_level = nullptr;
  pth_status SameLevelNeighbors();
  // patches of field f (built on first use); null with an error if f has no fluxes
  FluxCorrection *FluxPatches(int f);
  // adaptive refinement
  Adaptive amr;
  void SetRefinement(int b, int tag);          // MeshRefinement::SetRefinement
  pth_status Tag(MeshData &md);                // Refinement::Tag on md
  pth_status Adapt(int *modified);             // LoadBalancingAndAdaptiveMeshRefinement
  // takes the block list, storage and tables of nm (built for the new tree); nm receives the
  // old ones and frees them when destroyed
  void Adopt(Mesh &nm);
};

// Optional per-launch timing with HIP events on the launch stream (instrumentation regions
// like the reference's PARTHENON_INSTRUMENT, src/utils/instrument.hpp).
struct ProfRecord {
  const char *name;
  hipEvent_t start, stop;
};
struct Profiler {
  bool on = false;
  std::vector<ProfRecord> records;
  std::vector<hipEvent_t> pool;
  hipEvent_t Take();
  void Clear();
};
Profiler &Prof();

class ProfScope {  // records start/stop events around the launches in its lifetime
 public:
  ProfScope(const char *name, hipStream_t s);
  ~ProfScope();

 private:
  const char *name_;
  hipStream_t s_;
  hipEvent_t a_ = nullptr;
};

// kernels (pth_bvals.hip)
pth_status LaunchPack(const Table &t, double *arena_unused, hipStream_t s, bool to_buffer,
                      int mode);
bool CanExchangeDirect(const Table &pack, const Table &unpack);
pth_status LaunchExchangeDirect(const Table &pack, const Table &unpack, hipStream_t s);
// same-level channels of the listed bricks by ghost shells (cubic blocks of 8, 16 or 32 cells
// with 2 to 4 ghosts: ShellSupported)
bool ShellSupported(const Topology &t);
pth_status LaunchShell(const Topology &t, const std::vector<ShellBrick> &bricks,
                       const int *same, const int *order, hipStream_t s);
pth_status LaunchRestrict(const Table &t, const DeviceBlockGeom *geom, hipStream_t s);
pth_status LaunchProlong(const Table &t, const DeviceBlockGeom *geom, hipStream_t s);
pth_status LaunchBoundaryConditions(const BcTable &t, hipStream_t s);
// box copies: region q copies its box from a (at lo, strides astride) to b (pre-offset to the
// destination box origin, strides bstride)
pth_status LaunchCopy(const Table &t, hipStream_t s);

std::uint64_t Magic(int d);

}  // namespace pth

inline pth::Mesh *Impl(pth_mesh *m) { return reinterpret_cast<pth::Mesh *>(m); }
inline const pth::Mesh *Impl(const pth_mesh *m) { return reinterpret_cast<const pth::Mesh *>(m); }
inline pth::MeshData *Impl(pth_meshdata *m) { return reinterpret_cast<pth::MeshData *>(m); }
