// Host launchers of the Burgers stage kernels (pth_burgers.hip).
#pragma once
#include "pth_internal.hpp"

namespace pth {

pth_status BurgersRecon(const Topology &t, int V, const double *U, double *face, int recon,
                        hipStream_t s);
pth_status BurgersFlux(const Topology &t, int V, const double *face, double *flux,
                       hipStream_t s);
pth_status FluxDivergence(const Topology &t, int V, const double *flux, double *dudt,
                          const BlockCoefs *coef, hipStream_t s);
pth_status WeightedSum(const double *x, const double *y, double w1, double w2, double *z,
                       std::int64_t n, hipStream_t s);
pth_status BurgersDerived(const Topology &t, int V, const double *U, double *out,
                          hipStream_t s);
pth_status BurgersDtMin(const Topology &t, int V, const double *U, const BlockCoefs *coef,
                        double *partial, int npartial, double *out, hipStream_t s);
// workgroup partials the CFL reduction uses on topology t (at most 8192)
int BurgersDtPartials(const Topology &t);
// out[0] = cfl * (the CFL minimum), device memory, asynchronous on s; partial holds npartial
// doubles and must cover BurgersDtPartials(t)
pth_status BurgersDtToDevice(const Topology &t, int V, const double *U, const BlockCoefs *coef,
                             double *partial, int npartial, double cfl, double *out,
                             hipStream_t s);

// fused stage path
pth_status BurgersReconTasks(const Topology &t, int V, const double *U, double *face, int recon,
                             hipStream_t s);
int FusedUpdatePartials(const Topology &t);
pth_status BurgersFusedUpdate(const Topology &t, int V, const double *face, const double *U0,
                              const double *base, double *out, double *derived, double *flux,
                              double *partial, const BlockCoefs *coef, double beta,
                              const double *dtvar, const FluxCorrection *fc, hipStream_t s);
// flux correction: restricted fine fluxes into the correction store, from the face states of
// the fused stage or from stored fluxes; and the store into the coarse blocks' stored fluxes
pth_status FlxcorFromFaces(const Topology &t, int V, const FluxCorrection &fc,
                           const double *face, hipStream_t s);
pth_status FlxcorSend(const Topology &t, const FluxCorrection &fc, const double *flux,
                      hipStream_t s);
pth_status FlxcorSet(const Topology &t, const FluxCorrection &fc, double *flux, hipStream_t s);
// plane-tiled reconstruction + HLL into the flux array, and the update from stored fluxes
bool ReconFluxSupported(const Topology &t, int V);
// part 0: every face; 1: faces independent of ghost cells; 2: the remaining faces
bool BurgersSweepSupported(const Topology &t, int V);
// one launch per stage: reconstruction, HLL, flux divergence, RK update, derived, CFL partial
// per workgroup (partial[0 .. *groups), at most max_partials workgroups); faces with a finer
// neighbour take fc's corrected fluxes; waves (3 or 4) is the register budget the kernel is
// compiled for (waves per SIMD); spread (1, 2 or 3) divides the resident workgroup count, so
// each workgroup takes longer plane segments (fewer z-state rebuilds on small meshes); with
// push (27 same-level neighbours per block, or -1) the output's boundary cells are also
// written into those neighbours' ghost cells of out; with list (nlist block ids, device
// memory; 16^3 blocks only) the launch covers those blocks alone
pth_status BurgersSweep(const Topology &t, int V, const double *U, const double *base,
                        double *out, double *derived, double *partial, int max_partials,
                        const BlockCoefs *coef, double beta, const double *dtvar, int recon,
                        const FluxCorrection *fc, int waves, int spread, const int *push,
                        int *groups, hipStream_t s, const int *list = nullptr, int nlist = 0);
// flux-correction send of the sweep: fine fluxes from the state, restricted into fc's store
pth_status FlxcorDirect(const Topology &t, int V, const FluxCorrection &fc, const double *U,
                        int recon, hipStream_t s);
pth_status BurgersReconFlux(const Topology &t, int V, const double *U, double *flux, int recon,
                            int part, int ppw, hipStream_t s);
pth_status BurgersFluxUpdate(const Topology &t, int V, const double *flux, const double *U0,
                             const double *base, double *out, double *derived, double *partial,
                             const BlockCoefs *coef, double beta, const double *dtvar,
                             hipStream_t s);
// device clock slots: {dt of the next cycle (0 once time >= tlim), time, CFL estimate,
// dt of the last cycle run, the driver's tm.dt}
constexpr int kClockSlots = 5;
pth_status BurgersClock(const double *partial, int n, double cfl, double *dtvar, int advance,
                        const ClockLimits &lim, hipStream_t s);
// octant sums of U^2 * volume / mesh volume (8 values into out, device memory)
pth_status BurgersMassHistory(const Topology &t, int V, const double *U,
                              const DeviceBlockGeom *geom, const BlockCoefs *coef,
                              const double mesh_lo[3], const double mesh_hi[3], double *scratch,
                              int nscratch, double *out, hipStream_t s);

pth_status LaunchRestrictT(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                           hipStream_t s);
pth_status LaunchProlongT(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                          hipStream_t s, bool minmod);

}  // namespace pth
