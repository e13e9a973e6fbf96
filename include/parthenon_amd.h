/*
 * parthenon_amd — C ABI of the MI355X-native per-stage meshblock data path.
 *
 * Plain pointers, sizes and opaque handles only. Every entry point returns a pth_status
 * mirroring the reference's TaskStatus (src/basic_types.hpp:59); on PTH_FAIL the reason
 * is available from pth_last_error(). Device work runs on the stream set with
 * pth_mesh_set_stream (default: the null stream) and is asynchronous unless stated.
 *
 * Reference interfaces each entry replaces are cited as file:line relative to the
 * reference repository root.
 */
#ifndef PARTHENON_AMD_H_
#define PARTHENON_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* TaskStatus {complete, incomplete, iterate, fail}: src/basic_types.hpp:59 */
typedef enum {
  PTH_COMPLETE = 0,
  PTH_INCOMPLETE = 1,
  PTH_ITERATE = 2,
  PTH_FAIL = 3
} pth_status;

/* BoundaryType: src/basic_types.hpp:156 (any, local, nonlocal, flxcor_send, flxcor_recv;
 * the reference's gmg_* types are not offered: the multigrid solver keeps its level transfers
 * inside pth_gmg_* and pth_mg_*). On one GPU every neighbour is local, so any == local and
 * nonlocal has no buffers. */
typedef enum {
  PTH_BT_ANY = 0,
  PTH_BT_LOCAL = 1,
  PTH_BT_NONLOCAL = 2,
  PTH_BT_FLXCOR_SEND = 3,
  PTH_BT_FLXCOR_RECV = 4
} pth_boundary_type;

/* Metadata flags used by the path: src/interface/metadata.hpp (FillGhost, Independent,
 * WithFluxes, OneCopy). Only cell-centred fields are supported. */
enum {
  PTH_FILL_GHOST = 1u << 0,
  PTH_INDEPENDENT = 1u << 1,
  PTH_WITH_FLUXES = 1u << 2,
  PTH_ONE_COPY = 1u << 3,
  PTH_VECTOR = 1u << 4 /* Metadata::Vector: component v is the x(v+1) vector component */
};

/* How SendBoundBufs/SetBounds launch: one persistent table-walking kernel per call (default),
 * or the plain form with one kernel launch per boundary buffer. */
typedef enum { PTH_PACK_PERSISTENT = 0, PTH_PACK_PER_BUFFER = 1 } pth_pack_mode;

/* Boundary condition of a domain face (<parthenon/mesh> ix1_bc ... ox3_bc, src/mesh/mesh.cpp:
 * 1180-1185; the reference's "outflow" and "reflecting" functions, src/bvals/
 * boundary_conditions.cpp:89-137; FixedFace with value 0, GenericBC<..., BCType::FixedFace>
 * of boundary_conditions_generic.hpp:229-232 as the poisson_gmg package registers it,
 * example/poisson_gmg/poisson_package.cpp:60-78: ghost = 2 * 0 - mirror cell, every
 * component). Faces are numbered like BoundaryFace: 0 inner x1, 1 outer x1, 2 inner x2,
 * 3 outer x2, 4 inner x3, 5 outer x3. */
typedef enum {
  PTH_BC_PERIODIC = 0,
  PTH_BC_OUTFLOW = 1,
  PTH_BC_REFLECT = 2,
  PTH_BC_FIXED_FACE = 3
} pth_bc;

typedef struct pth_mesh pth_mesh;         /* Mesh: blocks, neighbours, fields, comm buffers */
typedef struct pth_meshdata pth_meshdata; /* MeshData: one named container over all blocks */

/* Mesh construction. Replaces Mesh::Mesh + static refinement (src/mesh/mesh.cpp:1054-1116)
 * and Mesh::SetMeshBlockNeighbors (src/mesh/mesh-gmg.cpp:48-94). nx: mesh cells,
 * bnx: meshblock cells, periodic: per-dimension flag (non-periodic faces have no neighbour). */
pth_mesh *pth_mesh_create(const int nx[3], const int bnx[3], int nghost, const double xmin[3],
                          const double xmax[3], const int periodic[3]);
/* Condition of one face of a non-periodic dimension (default PTH_BC_OUTFLOW; the faces of a
 * periodic dimension stay periodic). Must precede pth_mesh_plan / pth_mesh_finalize. */
pth_status pth_mesh_set_boundary(pth_mesh *m, int face, int type);
/* <parthenon/static_refinementN> box {x1min,x1max,x2min,x2max,x3min,x3max} at level. */
pth_status pth_mesh_refine_box(pth_mesh *m, const double box[6], int level);
/* Union of refinement boxes, one per level-`level` block intersecting the sphere. */
pth_status pth_mesh_refine_sphere(pth_mesh *m, const double center[3], double radius, int level);
/* StateDescriptor::AddField (src/interface/state_descriptor.cpp) for a cell-centred field of
 * ncomp components. Must precede pth_mesh_finalize. */
pth_status pth_mesh_add_field(pth_mesh *m, const char *name, int ncomp, unsigned flags);
/* Metadata::RegisterRefinementOps<ProlongationOp, RestrictionOp> (src/interface/metadata.hpp:
 * 555) of one field, resolved like StateDescriptor::RefinementFncsToIDs (state_descriptor.hpp:
 * 256): the shared prolongation of its coarse buffers into fine ghost cells, and of refined
 * blocks' interiors on a remesh, is ProlongateSharedMinMod (the default) or
 * ProlongateSharedLinear (src/prolong_restrict/pr_ops.hpp:166-289); restriction is
 * RestrictAverage. Must precede pth_mesh_finalize. */
enum { PTH_PROLONG_MINMOD = 0, PTH_PROLONG_LINEAR = 1 };
enum { PTH_RESTRICT_AVERAGE = 0 };
pth_status pth_mesh_register_refinement_ops(pth_mesh *m, const char *field, int prolong,
                                            int restrict_op);
/* Builds blocks and neighbour links on the host only (no device work); idempotent.
 * pth_mesh_finalize calls it first. */
pth_status pth_mesh_plan(pth_mesh *m);
/* Builds blocks, neighbours, the device buffer tables and communication buffers. */
pth_status pth_mesh_finalize(pth_mesh *m);
void pth_mesh_destroy(pth_mesh *m);
pth_status pth_mesh_set_stream(pth_mesh *m, void *hip_stream);
pth_status pth_mesh_synchronize(pth_mesh *m);
pth_status pth_mesh_set_pack_mode(pth_mesh *m, int mode);

int pth_mesh_nblocks(const pth_mesh *m);
int pth_mesh_multilevel(const pth_mesh *m);
/* out: level, lx1, lx2, lx3 ; x: xmin[3], xmax[3] of block gid. */
pth_status pth_mesh_block_info(const pth_mesh *m, int gid, int64_t out[4], double x[6]);
/* cellbounds n[3], ng[3] then c_cellbounds n[3], ng[3] (src/mesh/meshblock.cpp:201-218) */
pth_status pth_mesh_index_shapes(const pth_mesh *m, int out[12]);
/* Neighbour links of block gid (reference NeighborBlock): up to max records of 8 int64
 * {gid, level, lx1, lx2, lx3 (as seen from gid, not wrapped), ox1, ox2, ox3}; returns count. */
int pth_mesh_neighbors(const pth_mesh *m, int gid, int64_t *out, int max);
/* Index boxes of every boundary region of block gid, one record of 7 ints per neighbour link
 * and kind {kind, i_lo, i_hi, j_lo, j_hi, k_lo, k_hi}. kind: 0 send, 1 receive, 2 restriction
 * before send, 3 prolongation after set, 4 restriction after set (kinds 2-4 in coarse index
 * space), 5 flux-correction restriction before send (face fluxes, coarse index space), 6
 * flux-correction receive (loop_utils.hpp:146-160, bnd_info.cpp:387-403). Returns the number
 * of records written (kinds absent for a link are skipped). */
int pth_mesh_region_boxes(const pth_mesh *m, int gid, int *out, int max);
/* Boundary table sizes for bound type bt: number of buffers and of buffer elements used. */
pth_status pth_mesh_exchange_counts(const pth_mesh *m, int bt, int64_t *nbuffers,
                                    int64_t *nelements);

/* MeshData<Real> lookup/creation by name (src/interface/mesh_data.hpp:191; the
 * mesh_data.GetOrAdd calls of burgers_driver.cpp:76-79). Storage for non-OneCopy fields is
 * allocated on first use; OneCopy fields are shared by every container. */
pth_meshdata *pth_meshdata_get(pth_mesh *m, const char *name);

/* Device pointers of a field: data [nblocks][ncomp][Ek][Ej][Ei]; coarse (multilevel) in the
 * c_cellbounds shape; flux [3][nblocks][ncomp][Ek][Ej][Ei] (x-face i at index i). */
pth_status pth_field_pointers(pth_meshdata *md, const char *field, double **data,
                              double **coarse, double **flux);
/* Synchronous host<->device copies of a whole field (n = number of doubles). which: 0 data,
 * 1 coarse, 2 flux. */
pth_status pth_field_upload(pth_meshdata *md, const char *field, int which, const double *host,
                            size_t n);
pth_status pth_field_download(pth_meshdata *md, const char *field, int which, double *host,
                              size_t n);

/* ---- packs: SparsePack / VariablePack / VariableFluxPack over a MeshData ----
 * PDOpt (src/interface/sparse_pack_base.hpp:46): with fluxes, coarse buffers, flattened
 * (block, variable) index. */
enum { PTH_PACK_WITH_FLUXES = 1u << 0, PTH_PACK_COARSE = 1u << 1, PTH_PACK_FLATTEN = 1u << 2 };

/* Device view of one pack. Every pointer is device memory owned by the library and valid
 * until the mesh is remeshed. Pack index v runs over the requested fields in order, a field
 * of ncomp components taking ncomp consecutive indices; every block holds every field (no
 * sparse fields on this path), so bounds are the same for all blocks. */
typedef struct {
  double *const *data;   /* [nblocks * nvar]: brick of (block, v), [k][j][i] over ent[] */
  double *const *flux;   /* [3 * nblocks * nvar]: face-flux bricks of direction dir - 1, or NULL */
  const int *bounds;     /* [2 * nfields]: first pack index of field f at f, last at nfields + f */
  const double *coords;  /* [nblocks * 8]: dx1, dx2, dx3, area1, area2, area3, volume, 0 */
  const double *geom;    /* [nblocks * 12]: x0[3], dx[3], coarse x0[3], coarse dx[3]; cell
                            centre of index i along d is x0[d] + (i + 0.5) dx[d] */
  int nblocks, nvar, nfields, options;
  int ent[3];            /* brick extents (i, j, k); the coarse extents with PTH_PACK_COARSE */
  int is, ie, js, je, ks, ke; /* interior index range (inclusive) of those bricks */
} pth_pack_view;

typedef struct pth_pack pth_pack;
/* MakePackDescriptor(md, vars, {}, options).GetPack(md) and MeshData::PackVariables(
 * AndFluxes)(vars) (src/interface/make_pack_descriptor.hpp:42-105, sparse_pack.hpp:53-409,
 * sparse_pack_base.cpp:95-346, mesh_data.hpp:349-430): the pack of the named fields, built on
 * first use and cached in md by (fields, options); after a remesh the next call rebuilds it.
 * Fails (NULL) for an unknown field, for fluxes of a field without them, for coarse buffers
 * on a single-level mesh. */
pth_pack *pth_pack_get(pth_meshdata *md, const char *const *fields, int nfields,
                       unsigned options);
pth_status pth_pack_view_get(const pth_pack *p, pth_pack_view *view);
/* GetLowerBoundHost / GetUpperBoundHost (sparse_pack.hpp:164-190) of requested field f. */
pth_status pth_pack_bounds(const pth_pack *p, int field, int *lo, int *hi);

/* ---- boundary communication (src/bvals/comms/bvals_in_one.hpp:41-93) ---- */
/* SendBoundBufs<bt>: boundary_communication.cpp:49-159 (restrict + pack). */
pth_status pth_send_bound_bufs(pth_meshdata *md, int bt);
/* StartReceiveBoundBufs<bt>: boundary_communication.cpp:173-186. */
pth_status pth_start_receive_bound_bufs(pth_meshdata *md, int bt);
/* ReceiveBoundBufs<bt>: boundary_communication.cpp:201-236 (state flip on one GPU). */
pth_status pth_receive_bound_bufs(pth_meshdata *md, int bt);
/* SetBounds<bt>: boundary_communication.cpp:251-348 (unpack + restrict). */
pth_status pth_set_bounds(pth_meshdata *md, int bt);
/* ProlongateBounds<bt>: boundary_communication.cpp:361-393 (minmod shared prolongation). */
pth_status pth_prolongate_bounds(pth_meshdata *md, int bt);
/* ApplyBoundaryConditionsOnCoarseOrFineMD (src/bvals/boundary_conditions.cpp:82-87): the
 * physical boundary conditions of every FillGhost field on the blocks' data (coarse = 0) or
 * coarse buffers (coarse = 1), GenericBC Outflow / Reflect (boundary_conditions_generic.hpp:
 * 174-255) on each face that lies on the domain boundary of a non-periodic dimension. The
 * faces are applied in one launch whose result equals the reference's face-by-face order. */
pth_status pth_apply_boundary_conditions(pth_meshdata *md, int coarse);

/* Send + Receive + Set (+ coarse-buffer boundary conditions and ProlongateBounds on multilevel
 * meshes, then the boundary conditions: the order of AddBoundaryExchangeTasks,
 * boundary_communication.cpp:432-445) of bound type any in one call. On one GPU the buffer stage is skipped: every element moves from its sending cell to
 * its receiving cell through the same region tables in one launch (send regions read only
 * interior cells, receive regions write only ghost cells). Ghost data is identical to the
 * buffered sequence. No counterpart in the reference, whose exchange is split into tasks so
 * that MPI messages can overlap; it is for single-GPU drivers that do not need the buffers. */
pth_status pth_exchange_fused(pth_meshdata *md, int bt);

/* ---- Burgers benchmark package and Update:: tasks ---- */
/* burgers_package::CalculateFluxes (benchmarks/burgers/burgers_package.cpp:202-404) on the
 * field "U" of md; recon 0 = WENO5Z, 1 = MC-limited linear. The reconstructed face states
 * (the reference's OneCopy Ulx..Urz) live in a library-owned scratch arena. */
pth_status pth_burgers_calculate_fluxes(pth_meshdata *md, int recon);
/* Update::FluxDivergence<MeshData> (src/interface/update.cpp:63-84). */
pth_status pth_flux_divergence(pth_meshdata *in, pth_meshdata *dudt);
/* Update::WeightedSumData over Independent fields (src/interface/update.hpp:71-90):
 * out = w1 * in1 + w2 * in2 over entire arrays. */
pth_status pth_weighted_sum_independent(pth_meshdata *in1, pth_meshdata *in2, double w1,
                                        double w2, pth_meshdata *out);
/* burgers_package::CalculateDerived (burgers_package.cpp:143-167). */
pth_status pth_burgers_calculate_derived(pth_meshdata *md);
/* burgers_package::EstimateTimestepMesh (burgers_package.cpp:170-200); synchronous. */
pth_status pth_burgers_estimate_timestep(pth_meshdata *md, double cfl, double *dt);
/* The same estimate without allocation or host synchronization; no counterpart in the
 * reference. For callers that cannot allocate device or pinned memory (code under hipGraph
 * capture, tight serving loops): asynchronous on the mesh's stream, it writes cfl * min into
 * dt_device[0] (device memory, the same bits as pth_burgers_estimate_timestep) and keeps the
 * reduction's workgroup partials in a static 64 KiB device array per GPU (at most 8192
 * partials; a larger need is rejected). Calls that overlap on one GPU share that array and
 * race: order them on one stream. It calls no hipMalloc, hipFree or allocating routine. */
pth_status pth_burgers_estimate_timestep_noalloc(pth_meshdata *md, double cfl, double *dt_device);

/* ---- device-resident Burgers driver (the task order of BurgersDriver::MakeTaskCollection,
 * benchmarks/burgers/burgers_driver.cpp:53-148, RK2 of low_storage_integrator.cpp:67-83 and
 * EvolutionDriver::SetGlobalTimeStep, src/driver/driver.cpp:210-235) on the containers "base"
 * and "1". Each stage is one reconstruction launch, one fused flux/divergence/update launch
 * and the boundary exchange; dt and time stay in device memory, so a cycle needs no host
 * synchronization and can be captured once in a hipGraph and replayed. Single-level meshes. */
/* Mesh::Initialize after the problem generator on "base": exchange, FillDerived, first dt. */
pth_status pth_burgers_init(pth_mesh *m, double cfl, int recon);
/* ncycles RK2 cycles; use_graph != 0 captures one cycle into a hipGraph on first use and
 * replays it. Asynchronous on the driver's own stream. */
pth_status pth_burgers_run_cycles(pth_mesh *m, int ncycles, int use_graph);
/* Driver stage form (no counterpart in the reference). The two plane forms give the same
 * bits; the block sweep is built with FMA contraction and agrees with them to rounding (about
 * 1e-15 relative after a few cycles). For cubic blocks of 8 or 16 cells with 4 ghosts a stage is either one block-sweep launch
 * (reconstruction, HLL, flux divergence and RK update per block, no face state or flux in
 * HBM) followed by the exchange, or the plane-tiled reconstruction into the flux array and
 * the update, with the stage input's exchange either beside the reconstruction of the faces
 * whose stencils read no ghost cell (on a second stream) or after the update. Unless one of
 * the setters below is called first, pth_burgers_init times whole cycles of each form on the
 * mesh's state (restored afterwards) and keeps the fastest. set_sweep(on != 0) selects the
 * sweep; set_overlap chooses between the two plane forms when the sweep is off.
 * pth_burgers_form reports the form (PTH_INCOMPLETE before the automatic choice). */
pth_status pth_burgers_set_overlap(pth_mesh *m, int on);
pth_status pth_burgers_set_sweep(pth_mesh *m, int on);
pth_status pth_burgers_form(pth_mesh *m, int *sweep, int *overlap);
/* With on != 0 the automatic choice of pth_burgers_init considers only the bit-exact plane
 * forms (when the mesh supports them), so a run's bits cannot depend on timing noise; must
 * precede pth_burgers_init. No counterpart in the reference. */
pth_status pth_burgers_set_bit_exact(pth_mesh *m, int on);
/* Sweep form on statically refined meshes of 16^3 blocks with 4 ghosts (on by default; no
 * counterpart in the reference): each stage sweeps the blocks without a neighbour at another
 * level or a physical face first, while the cross-level part of the stage input's exchange
 * (restriction, channels between levels, prolongation, conditions) runs on a second stream,
 * then the flux-correction send and the sweep of the other blocks; the last cycle's
 * cross-level exchange runs at the end of pth_burgers_run_cycles. The same bits as on = 0. */
pth_status pth_burgers_set_sweep_split(pth_mesh *m, int on);
/* Driver tuning (no counterpart in the reference): planes per workgroup (1, 2 or 4) of the
 * plane-tiled reconstruction for its three forms (all faces, ghost-free faces, the rest), and
 * the block sweep's register budget (the waves per SIMD it is compiled for: 2 or 3 for 16^3
 * blocks with 4 ghosts, 3 or 4 otherwise; other values are clamped) as planes[3] of
 * pth_burgers_tuning. pth_burgers_init picks them by timing each choice on the mesh's data
 * unless set here first; pth_burgers_tuning returns PTH_INCOMPLETE while no choice has been
 * made. Results do not depend on the choice. */
pth_status pth_burgers_tuning(pth_mesh *m, int planes[4]);
pth_status pth_burgers_set_tuning(pth_mesh *m, const int planes[4]);
/* Synchronous read of the device clock: the driver's dt (tm.dt) and simulated time. */
pth_status pth_burgers_clock(pth_mesh *m, double *dt, double *time);
/* <parthenon/time> limits of EvolutionDriver (src/driver/driver.hpp:66-94; applied as in
 * SetGlobalTimeStep, driver.cpp:210-269): lim = {tlim, dt_init, dt_user, dt_force, dt_floor,
 * dt_ceil}; the reference defaults are {inf, DBL_MAX, DBL_MAX, -DBL_MAX, DBL_MIN, DBL_MAX},
 * dt_init_force 0. The last cycle before tlim is shortened to end on it. The device cannot
 * end the host's loop, so a cycle run after time >= tlim (where SimTime::KeepGoing,
 * src/basic_types.hpp:241, would have stopped the reference) advances nothing: the state,
 * time and dt stay as they are. Takes effect from the next pth_burgers_init or cycle. */
pth_status pth_burgers_set_time_limits(pth_mesh *m, const double lim[6], int dt_init_force);
/* Waits for the driver's stream. */
pth_status pth_burgers_synchronize(pth_mesh *m);

/* ---- adaptive mesh refinement ----
 * <parthenon/mesh> refinement = adaptive with numlevel and derefine_count (mesh.cpp:179,
 * mesh_refinement.cpp:55-58). Must precede pth_mesh_plan / pth_mesh_finalize; the mesh keeps
 * coarse buffers on every level. Levels here count from the base grid; the reference's
 * logical level is root_level + level, root_level = log2 of the largest power of two dividing
 * the base block count of every dimension (forest.cpp:64-105). */
pth_status pth_mesh_set_adaptive(pth_mesh *m, int numlevel, int derefine_count);
/* One <parthenon/refinementN> block (AMRCriteria, amr_criteria.cpp:27-68, 88-106): method 1
 * derivative_order_1, 2 derivative_order_2, on component comp (vector_i) of field; max_level
 * <= 0 means numlevel and larger values are clamped to it. As in the reference, max_level is
 * compared with the logical level (refinement_package.cpp:78). */
pth_status pth_mesh_add_refinement_criterion(pth_mesh *m, int method, const char *field, int comp,
                                             double refine_tol, double derefine_tol,
                                             int max_level);
/* Refinement::Tag<MeshData> (refinement_package.cpp:153-173): every criterion on md's data in
 * one launch per criterion, then CheckAllRefinement and MeshRefinement::SetRefinement
 * (mesh_refinement.cpp:85-117) per block. Synchronous (the flags live on the host, like the
 * reference's refine_flag_). Ghost cells of the tagged fields must be current. */
pth_status pth_refinement_tag(pth_meshdata *md);
/* MeshRefinement::SetRefinement(AmrTag) of one block (tag -1, 0 or 1). */
pth_status pth_mesh_set_refinement(pth_mesh *m, int gid, int tag);
/* refine_flag_ and deref_count_ of every block (arrays of pth_mesh_nblocks entries). */
pth_status pth_mesh_refinement_flags(const pth_mesh *m, int *flags, int *counts);
int pth_mesh_root_level(const pth_mesh *m);
/* Mesh::LoadBalancingAndAdaptiveMeshRefinement on one GPU (mesh-amr_loadbalance.cpp:324-351):
 * UpdateMeshBlockTree (:496-626) from the flags; if the tree changed, the block list, every
 * MeshData's fields, the region tables and the buffers are rebuilt in place and the data moved
 * (RedistributeAndRefineMeshBlocks, :663-1003): restriction of derefined blocks, copies of
 * moved blocks and octants, interior prolongation of refined blocks, then the boundary
 * exchange of "base" (CommunicateBoundaries). MeshData handles stay valid; data pointers and
 * block ids change. *modified = 1 when the mesh changed. */
pth_status pth_mesh_adapt(pth_mesh *m, int *modified);
/* EvolutionDriver::Execute after a modified mesh (driver.cpp:128-131): FillDerived,
 * InitializeBlockTimeSteps on the new blocks and SetGlobalTimeStep from the dt of the last
 * cycle, on the device clock of pth_burgers_init. When the block count changed by a factor of
 * two or more since the stage form and tuning were timed, and the caller fixed neither, both
 * are timed again on the new blocks (the state is restored). */
pth_status pth_burgers_after_remesh(pth_mesh *m);

/* History: the 8 octant masses sum(U^2 * cell volume / mesh volume) of the package's history
 * variables "MS Mass 0..7" (burgers_package.cpp:96-135, MassHistory :406-439), octants in the
 * reference order (x outer, then y, then z); synchronous. */
pth_status pth_burgers_mass_history(pth_meshdata *md, double out[8]);

/* Problem generators for "U" of md (host-evaluated, uploaded): the seeded sum of sinusoids
 * and the reference's tanh/exp profile (benchmarks/burgers/parthenon_app_inputs.cpp:34-79). */
pth_status pth_burgers_pgen_sinusoid(pth_meshdata *md, uint64_t seed);
pth_status pth_burgers_pgen_vibe(pth_meshdata *md);

/* ---- geometric-multigrid Poisson solver ----
 * One solver object per N^3 cell-centred problem on [0,1]^3 with FixedFace-0 boundaries:
 * the reference's MGSolver<u, rhs, PoissonEquation> (src/solvers/mg_solver.hpp:37-521,
 * example/poisson_gmg/poisson_equation.hpp:41-205, poisson_package.cpp:60-130) on the uniform
 * GMG hierarchy of src/mesh/mesh-gmg.cpp:103-154 (levels halve while the count is even).
 * precision 0: V-cycles in FP32 inside an FP64 iterative-refinement loop (residual and update
 * in FP64); precision 1: the same solver in FP64. omega, pre and post: weighted-Jacobi weight
 * and sweeps (the reference's MGParams smoother, mg_solver.hpp:312-345, fixes SRJ weights).
 * Arrays are the N^3 interior cells in [k][j][i] order (host memory). */
typedef struct pth_gmg pth_gmg;
pth_gmg *pth_gmg_create(int n, int precision, double omega, int pre, int post);
void pth_gmg_destroy(pth_gmg *g);
int pth_gmg_levels(const pth_gmg *g);
/* RHS: 16 Gaussian blobs s exp(-|x-c|^2/w^2), c ~ U(0.1,0.9)^3, w ~ U(0.02,0.08), s = +-1,
 * from SplitMix64(seed) in blob order (cx, cy, cz, w, s); resets the solution to 0. */
pth_status pth_gmg_set_rhs_blobs(pth_gmg *g, uint64_t seed);
pth_status pth_gmg_set_rhs(pth_gmg *g, const double *host, size_t n);
pth_status pth_gmg_get_rhs(pth_gmg *g, double *host, size_t n);
pth_status pth_gmg_get_solution(pth_gmg *g, double *host, size_t n);
pth_status pth_gmg_reset(pth_gmg *g);
/* V-cycles per FP64 refinement step (default 1); cycles after the first continue from the
 * current correction on the same right-hand side. */
pth_status pth_gmg_set_inner_cycles(pth_gmg *g, int k);
/* Tuning (no counterpart in the reference): with on != 0 (the default) Jacobi sweeps of levels
 * whose size is a multiple of 64 run by z columns (z neighbours kept in registers); the
 * results are the same bits either way. */
pth_status pth_gmg_set_column_sweeps(pth_gmg *h, int on);
/* MGSolver::AddTasks iteration (mg_solver.hpp:119-161): V-cycles until the relative L2
 * residual ||b - A u|| / ||b|| < tol or max_iters cycles; iters and final_rel reported. */
pth_status pth_gmg_solve(pth_gmg *g, double tol, int max_iters, int *iters, double *final_rel);
/* relative residual before every V-cycle and after the last; returns the count */
int pth_gmg_history(const pth_gmg *g, double *out, int max);

/* ---- solver field utilities over MeshData (src/solvers/solver_utils.hpp:151-277) ----
 * Cell fields of md with equal component counts; asynchronous on the mesh stream except the
 * dot product, which returns its value.
 * copy_data: CopyData<in, out> (:151-178), every cell (ghosts included), every component.
 * add_fields_and_store: AddFieldsAndStore<a, b, out> (:180-222), out = wa a + wb b on every
 *   cell (ghosts included); out may be a or b.
 * set_to_zero: SetToZero<var> (:224-250), interior cells.
 * dot_product: DotProductLocal<a, b> (:252-277) over every block of md, the sum over interior
 *   cells and components of a * b (fixed summation order: the same bits on every call). */
pth_status pth_solver_copy_data(pth_meshdata *md, const char *in, const char *out);
pth_status pth_solver_add_fields_and_store(pth_meshdata *md, const char *a, const char *b,
                                           const char *out, double wa, double wb);
pth_status pth_solver_set_to_zero(pth_meshdata *md, const char *var);
pth_status pth_solver_dot_product(pth_meshdata *md, const char *a, const char *b, double *out);

/* ---- the Poisson equation over MeshData (example/poisson_gmg/poisson_equation.hpp:41-205) ----
 * div(D grad u) - diagonal_alpha u = rhs with the package's diagonal_alpha (:83-84) and D either
 * constant or the reference's face field D (Metadata::Face, poisson_package.cpp:112-118):
 * D_field names a 3-component cell field whose component d at cell (i, j, k) holds D on the
 * cell's lower face along dimension d (the reference's pack(b, TE::F(d+1), D(), k, j, i)); a
 * block's upper faces along d sit in its first ghost cell along d, so fill D over the whole
 * brick (the problem generator's IndexDomain::entire, parthenon_app_inputs.cpp:44-78). */
typedef struct pth_poisson_equation {
  double D;              /* constant D (D_field null) */
  double diagonal_alpha;
  const char *D_field;   /* null, or the face field D */
} pth_poisson_equation;
/* PoissonEquation::Ax<x, out> (:41-53): CalculateFluxes (:131-184) then FluxMultiplyMatrix
 * (:186-205) on the interior cells of every block of md, without flux correction:
 * out = -alpha x + sum_d (F_d(lo) - F_d(hi)) / dx_d, F_d = D_d(face) (x(lo side) - x) / dx_d,
 * from x's cells and one ghost layer (fill them first: the boundary exchange with its
 * conditions). x and out: distinct one-component fields. The x3 face flux is the 7-point operator's (the reference stores the top x3 flux in
 * the x2 slot, :153-157). */
pth_status pth_poisson_ax(pth_meshdata *md, const pth_poisson_equation *eq, const char *x,
                          const char *out);

/* ---- MGSolver<u, rhs, PoissonEquation> over MeshData (src/solvers/mg_solver.hpp:37-520) ----
 * u (FillGhost) and rhs: one-component fields; D > 0 (constant, or the face field D) and
 * diagonal_alpha >= 0.
 * Two hierarchies, both stopping on the RMS residual of the MeshData problem:
 * - uniform single-level 3-D meshes of N^3 cells on a cube with fixed_face (FixedFace) on all
 *   six faces and smoother 0: whole-level arrays (cell counts halving while even, as
 *   src/mesh/mesh-gmg.cpp:103-154 gives on such a mesh), FP32 V-cycles inside FP64 iterative
 *   refinement (precision 0) or FP64 (precision 1); the problem is solved on the unit cube
 *   with rhs scaled by L^2 / D (exact when L = D = 1), diagonal_alpha = 0;
 * - every other mesh or equation (static refinement, 1-D / 2-D, periodic faces, a face field
 *   D, diagonal_alpha != 0, or smoother 1-3): a hierarchy of block meshes (the leaf mesh, the
 *   mesh capped at each coarser refinement level, then the forest's merged and cell-halved
 *   levels below the base grid) in FP64 (precision 1 only); each level's ghosts come from
 *   its own boundary exchange with the faces' conditions, restriction is RestrictAverage and
 *   prolongation of the correction ProlongateSharedLinear; a face field D is read at pth_mg_setup and restricted down the
 *   levels by face averages (D is GMGRestrict in the reference, AddSetupTasks restricts it,
 *   mg_solver.hpp:188-203). u's registered prolongation operator defines the leaf problem's
 *   fine-coarse ghosts (the poisson_gmg package registers ProlongateSharedLinear). */
typedef struct pth_mg_params {
  int max_iters;             /* MGParams::max_iters (mg_solver.hpp:37-58): V-cycles at most */
  double residual_tolerance; /* stop when the RMS residual sqrt(sum r^2 / ncells) of
                                rhs - A u is below it (mg_solver.hpp:147-158) */
  int precision;             /* 0: FP32 V-cycles inside FP64 iterative refinement; 1: FP64 */
  int inner_cycles;          /* FP32 V-cycles per FP64 refinement step (precision 0), >= 1 */
  double omega;              /* weighted-Jacobi weight (smoother 0) */
  int pre_sweeps, post_sweeps; /* weighted-Jacobi sweeps (smoother 0) */
  int smoother;              /* 0: weighted Jacobi (omega, pre_sweeps, post_sweeps); 1, 2, 3:
                                MGParams::smoother "SRJ1" / "SRJ2" / "SRJ3", the reference's
                                stage weights by dimension (mg_solver.hpp:312-345), that many
                                stages before and after the coarser level */
} pth_mg_params;
typedef struct pth_mg pth_mg;
pth_mg *pth_mg_create(pth_mesh *m, const char *u, const char *rhs, const pth_mg_params *p,
                      const pth_poisson_equation *eq);
void pth_mg_destroy(pth_mg *mg);
/* MGSolver::AddSetupTasks (mg_solver.hpp:188-203): gathers rhs and the initial u of md's
 * blocks into the solver's FP64 finest level (after the work queued on the mesh stream) */
pth_status pth_mg_setup(pth_mg *mg, pth_meshdata *md);
/* MGSolver::AddTasks (mg_solver.hpp:119-161): at least one V-cycle, then V-cycles until the
 * RMS residual is below the tolerance or max_iters cycles ran; writes u's interior cells and
 * refills its ghosts (the boundary exchange of md, with the FixedFace conditions) */
pth_status pth_mg_solve(pth_mg *mg, pth_meshdata *md);
/* GetFinalIterations / GetFinalResidual (mg_solver.hpp:207-208): V-cycles run, RMS residual */
pth_status pth_mg_result(const pth_mg *mg, int *iterations, double *rms_residual);
/* levels of the solver's hierarchy, finest first (GetGMGMaxLevel - GetGMGMinLevel + 1) */
int pth_mg_levels(const pth_mg *mg);
/* RMS residual before every V-cycle and after the last; returns the count */
int pth_mg_history(const pth_mg *mg, double *rms, int max);
/* Launch form of the block-mesh hierarchy (meshes other than the uniform cube; no counterpart
 * in the reference, whose V-cycle is a task list run on the host): on (the default) the first
 * V-cycle of the solver runs eagerly and every later one replays one captured hipGraph of
 * the same launches, with the same results bit for bit; off: every V-cycle eagerly. The
 * whole-level solver of the uniform cube is unaffected. */
pth_status pth_mg_set_graph(pth_mg *mg, int on);

/* Serving loop of independent Poisson solves (no counterpart in the reference, which holds
 * one Mesh per process, src/parthenon_manager.cpp:170-173). Requests arrive at given times;
 * they are admitted in arrival order while a slot is free (max_active) and the device's free
 * HBM (hipMemGetInfo) holds the request's hierarchy plus a reserve (default 1 GiB); a
 * budget_bytes > 0 also caps the bytes the server holds (<= 0: no cap). Pooled hierarchies of
 * other sizes are freed first when memory is short. Every outer iteration runs the same
 * V-cycle stage of all active solves as one launch. */
typedef struct pth_gmg_server pth_gmg_server;
pth_gmg_server *pth_gmg_server_create(int precision, double omega, int pre, int post, double tol,
                                      int max_iters, int64_t budget_bytes, int max_active,
                                      int keep_solutions);
void pth_gmg_server_destroy(pth_gmg_server *s);
pth_status pth_gmg_server_submit(pth_gmg_server *s, int n, uint64_t seed, double arrival_s,
                                 int *id);
pth_status pth_gmg_server_set_inner_cycles(pth_gmg_server *s, int k);
pth_status pth_gmg_server_set_column_sweeps(pth_gmg_server *h, int on);
/* Free HBM (bytes) that admissions leave unallocated (default 1 GiB). */
pth_status pth_gmg_server_set_hbm_reserve(pth_gmg_server *s, int64_t bytes);
/* serve every submitted request; realtime: honour arrival times on the wall clock */
pth_status pth_gmg_server_run(pth_gmg_server *s, int realtime);
/* out: arrival, admission, completion (s, server clock), latency (s), final relative residual */
pth_status pth_gmg_server_result(const pth_gmg_server *s, int id, double out[5], int *iters);
pth_status pth_gmg_server_solution(const pth_gmg_server *s, int id, double *host, size_t n);
/* out: outer iterations run, largest batch, device bytes held; busy_s: time in iterations */
pth_status pth_gmg_server_stats(const pth_gmg_server *s, int64_t out[3], double *busy_s);

/* Instrumentation (the reference's PARTHENON_INSTRUMENT regions, src/utils/instrument.hpp):
 * when enabled, every kernel launch is bracketed by HIP events on its stream.
 * pth_profile_report waits for the recorded events, writes one line per kernel name,
 * "name launches total_ms\n", into buf (NUL-terminated, at most len bytes), clears the
 * records and returns the number of lines. No counterpart in the reference's FFI. */
pth_status pth_profile_enable(int enable);
int pth_profile_report(char *buf, int len);

/* Free and total memory of the current device (hipMemGetInfo); no counterpart in the reference. */
pth_status pth_device_memory(int64_t *free_bytes, int64_t *total_bytes);

const char *pth_last_error(void);
const char *pth_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PARTHENON_AMD_H_ */
