// Boundary-buffer kernels for gfx950: pack (SendBoundBufs), unpack (SetBounds), cell-average
// restriction and minmod / linear shared prolongation.
//
// Every call is ONE persistent launch walking a device table of boundary regions of uneven
// sizes (g-wide x-face strips, n-wide y/z strips, g^3 corners, coarse-fine regions). The host
// cuts every region into work items of at most kItem elements; a workgroup takes one item at a
// time, so the item and its region record are wavefront-uniform (scalar loads) and each lane
// decodes its own elements with multiply-shift divisions. Lanes of a wavefront touch
// consecutive buffer elements (fully coalesced 64-lane accesses of the buffer arena) and runs
// of consecutive i in the field brick. Each lane issues all its loads before its stores.
//
// Arithmetic follows the reference operators term by term (no FMA contraction; the library is
// built with -ffp-contract=off), so restriction and prolongation are reproducible bit for bit:
//   RestrictAverage         src/prolong_restrict/pr_ops.hpp:105-164
//   ProlongateSharedGeneral src/prolong_restrict/pr_ops.hpp:166-289, GradMinMod :95-101
//   pack / unpack loops     src/bvals/comms/boundary_communication.cpp:95-139, 273-334
#include <hip/hip_runtime.h>

#include "pth_device.hpp"
#include "pth_internal.hpp"

namespace pth {

namespace {

constexpr int kWG = 256;  // threads per workgroup (4 wavefronts of 64)
constexpr int kPer = kItemElems / kWG;  // elements per lane per item (16 loads in flight)

// The region tables hold generic pointers; every array they address is device global
// memory, accessed through Glob (global loads and stores).
typedef __attribute__((address_space(1))) double gdouble;

struct Cell {
  int i, j, k, c;
};

// element e of a region -> (i, j, k, comp) offsets within its box; exact for e < 2^30 / n
__device__ __forceinline__ Cell Decode(const Region &R, std::uint32_t e) {
  Cell d;
  const std::uint64_t q0 = (static_cast<std::uint64_t>(e) * R.magic[0]) >> 40;
  d.i = static_cast<int>(e - q0 * R.n[0]);
  const std::uint64_t q1 = (q0 * R.magic[1]) >> 40;
  d.j = static_cast<int>(q0 - q1 * R.n[1]);
  const std::uint64_t q2 = (q1 * R.magic[2]) >> 40;
  d.k = static_cast<int>(q1 - q2 * R.n[2]);
  d.c = static_cast<int>(q2);
  return d;
}

__device__ __forceinline__ std::int64_t VarOffset(const Region &R, const Cell &d) {
  return static_cast<std::int64_t>(d.c) * R.astride[2] +
         static_cast<std::int64_t>(d.k + R.lo[2]) * R.astride[1] +
         static_cast<std::int64_t>(d.j + R.lo[1]) * R.astride[0] + (d.i + R.lo[0]);
}

// Element of lane slot u within an item: lanes past the item's end take its last element
// and load and store the same value at the same address as the lane that owns it, so no
// load or store sits under a per-lane branch. (gfx9 counts loads and stores on
// one in-order counter; a store under a per-lane branch makes the compiler wait for every
// earlier access before it, which serializes the stores.)
__device__ __forceinline__ int Lane(const Item &I, int u) {
  const int local = u * kWG + static_cast<int>(threadIdx.x);
  return local < I.count ? local : I.count - 1;
}
__device__ __forceinline__ int Slots(const Item &I) { return (I.count + kWG - 1) / kWG; }

// Straight-line bodies for NS slots: the NS loads of a lane, then its NS stores. The kernels
// below pick NS = Slots(I) with a workgroup-uniform switch, so a small item (a corner, an
// edge) runs only the slots it fills and no load or store sits under a branch.
template <int NS, bool TO_BUFFER>
__device__ __forceinline__ void PackSlots(const Region &R, const Item &I) {
  double v[NS];
  int off[NS], e[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    e[u] = I.begin + Lane(I, u);
    off[u] = static_cast<int>(VarOffset(R, Decode(R, static_cast<std::uint32_t>(e[u]))));
    v[u] = TO_BUFFER ? Glob(R.a)[off[u]] : Glob(R.b)[e[u]];
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    if (TO_BUFFER)
      Glob(R.b)[e[u]] = v[u];
    else
      Glob(R.a)[off[u]] = v[u];
  }
}

template <int NS>
__device__ __forceinline__ void CopySlots(const Region &R, const Item &I) {
  double v[NS];
  std::int64_t off[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const Cell d = Decode(R, static_cast<std::uint32_t>(I.begin + Lane(I, u)));
    v[u] = Glob(R.a)[VarOffset(R, d)];
    off[u] = static_cast<std::int64_t>(d.c) * R.bstride[2] +
             static_cast<std::int64_t>(d.k) * R.bstride[1] +
             static_cast<std::int64_t>(d.j) * R.bstride[0] + d.i;
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) Glob(R.b)[off[u]] = v[u];
}

template <int NS>
__device__ __forceinline__ void ExchangeSlots(const Region &S, const Region &D, const Item &I) {
  double v[NS];
  std::int64_t off[NS];
#pragma unroll
  for (int u = 0; u < NS; ++u) {
    const std::uint32_t e = static_cast<std::uint32_t>(I.begin + Lane(I, u));
    v[u] = Glob(S.a)[VarOffset(S, Decode(S, e))];
    off[u] = VarOffset(D, Decode(D, e));
  }
#pragma unroll
  for (int u = 0; u < NS; ++u) Glob(D.a)[off[u]] = v[u];
}

// slot counts 1, 2, 4, 8 and kPer (rounded up) keep the code small
#define PTH_SLOTS(CALL)             \
  switch (Slots(I)) {               \
    case 1: CALL(1); break;         \
    case 2: CALL(2); break;         \
    case 3:                         \
    case 4: CALL(4); break;         \
    case 5:                         \
    case 6:                         \
    case 7:                         \
    case 8: CALL(8); break;         \
    default: CALL(kPer); break;     \
  }

// Pack (TO_BUFFER) or unpack: one launch over the whole table, one workgroup per work item.
// The item and its region record are uniform (scalar loads); each lane issues its loads
// before its stores (up to kPer of each), so a workgroup keeps up to 32 KiB in flight.
template <bool TO_BUFFER>
__global__ void __launch_bounds__(kWG)
    PackKernel(const Region *__restrict__ reg, const Item *__restrict__ items) {
  const Item I = items[blockIdx.x];
  const Region &R = reg[I.region];
#define PTH_PACK(NS) PackSlots<NS, TO_BUFFER>(R, I)
  PTH_SLOTS(PTH_PACK)
#undef PTH_PACK
}

// Box copies (remesh): region element e moves from its box in a to the same box position in
// b, which points at the destination box origin. Loads before stores, like the pack.
__global__ void __launch_bounds__(kWG)
    CopyKernel(const Region *__restrict__ reg, const Item *__restrict__ items) {
  const Item I = items[blockIdx.x];
  const Region &R = reg[I.region];
#define PTH_COPY(NS) CopySlots<NS>(R, I)
  PTH_SLOTS(PTH_COPY)
#undef PTH_COPY
}

// Fused exchange by region pairs: send region q and receive region q cover the same slice of
// the channel arena in the same element order, so element e of the pair moves from its cell
// in the send box to its cell in the receive box. Both records are uniform per work item; no
// per-element index is read. Loads before stores, like the pack.
__global__ void __launch_bounds__(kWG)
    ExchangeRegionKernel(const Region *__restrict__ sreg, const Region *__restrict__ rreg,
                         const Item *__restrict__ items) {
  const Item I = items[blockIdx.x];
  const Region &S = sreg[I.region];
  const Region &D = rreg[I.region];
#define PTH_XCHG(NS) ExchangeSlots<NS>(S, D, I)
  PTH_SLOTS(PTH_XCHG)
#undef PTH_XCHG
}

// Same-level ghost shells: one workgroup per (field brick q, receiving block b) fills every
// ghost cell of the brick whose neighbour across it is a block of the same level
// (same[27 b + dir] >= 0) from that neighbour's interior: cell (i, j, k) in the ghosts of
// direction (ox, oy, oz) is cell (i - ox N, j - oy N, k - oz N) of the neighbour's brick. The
// ghost cells are numbered in brick order, skipping the interior, so a wavefront writes
// consecutive addresses: whole planes below and above the interior, whole rows in the y ghost
// bands and, between them, the x ghost runs of consecutive rows (the high run of one row and
// the low run of the next are adjacent). Each lane issues its U loads before its U stores.
// Cells whose neighbour is at another level or across a physical face keep their value here
// and are filled by the cross-level channels and the boundary conditions after this launch.
template <int N, int G>
__global__ void __launch_bounds__(kWG) ShellKernel(const ShellArgs A) {
  constexpr int E = N + 2 * G, P = E * E, PI = P - N * N, S = 2 * G * P + N * PI;
  // loads in flight per lane: two passes over the shell where that takes at most 16
  constexpr int U = (S + 2 * kWG - 1) / (2 * kWG) < 16 ? (S + 2 * kWG - 1) / (2 * kWG) : 16;
  __shared__ int nbr[27];
  // logical id q * nblocks + m: each XCD takes a contiguous range of one brick's blocks, in
  // Morton order so that the neighbours reading a block's boundary run close in time
  const unsigned L = XcdRemap(blockIdx.x, gridDim.x);
  const int b = A.order[L % static_cast<unsigned>(A.nblocks)];
  const int q = static_cast<int>(L / static_cast<unsigned>(A.nblocks));
  if (threadIdx.x < 27) nbr[threadIdx.x] = A.same[27 * b + threadIdx.x];
  __syncthreads();
  const double *src0 = A.brick[q];
  const std::int64_t bs = A.stride[q];
  double *dst = A.brick[q] + b * bs;
  for (int e0 = static_cast<int>(threadIdx.x); e0 < S; e0 += U * kWG) {
    double v[U];
    int off[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = e0 + u * kWG < S ? e0 + u * kWG : S - 1;  // past the end: repeat the last
      {
        int i, j, k;
        if (e < G * P || e >= G * P + N * PI) {  // whole ghost planes
          const int t = e < G * P ? e : e - N * PI;
          k = t / P;
          k += k >= G ? N : 0;
          const int r = t % P;
          j = r / E;
          i = r % E;
        } else {
          const int t = e - G * P;
          k = G + t / PI;
          const int r = t % PI;
          if (r < G * E || r >= G * E + 2 * G * N) {  // whole ghost rows
            const int rr = r < G * E ? r : r - 2 * G * N;
            j = rr / E;
            j += j >= G ? N : 0;
            i = rr % E;
          } else {  // the x ghost runs of an interior row
            const int rr = r - G * E;
            j = G + rr / (2 * G);
            const int x = rr % (2 * G);
            i = x < G ? x : x + N;
          }
        }
        const int ox = (i >= G) + (i >= G + N) - 1, oy = (j >= G) + (j >= G + N) - 1,
                  oz = (k >= G) + (k >= G + N) - 1;
        const int sb = nbr[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)];
        off[u] = (k * E + j) * E + i;
        // a cell without a same-level neighbour keeps its value (it is rewritten here and
        // filled later by the cross-level channels or the boundary conditions), so every
        // load and store is unconditional
        v[u] = sb >= 0 ? src0[sb * bs + off[u] - (ox * N + oy * N * E + oz * N * P)]
                       : dst[off[u]];
      }
    }
    // streaming stores: the ghost lines are not read again by this launch, so they should
    // not push the neighbours' boundary lines out of the L2
#pragma unroll
    for (int u = 0; u < U; ++u) __builtin_nontemporal_store(v[u], &dst[off[u]]);
  }
}

// The plain form: one launch per region, one thread per element.
__global__ void __launch_bounds__(kWG) PackOneKernel(Region R, int to_buffer) {
  const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x;
  const std::int64_t count = static_cast<std::int64_t>(R.n[0]) * R.n[1] * R.n[2] * R.nc;
  if (e >= count) return;
  const std::int64_t off = VarOffset(R, Decode(R, static_cast<std::uint32_t>(e)));
  if (to_buffer)
    R.b[e] = R.a[off];
  else
    R.a[off] = R.b[e];
}

struct ProResParams {
  int flo[3];  // fine interior start per dimension
  int clo[3];  // coarse interior start per dimension
  int act[3];  // active dimensions
};

// Cell-average restriction of one coarse cell from its 2x2x2 fine children; the summation
// order is the reference's fixed pairwise order. Split into the loads (RestrictLoad: the
// eight children, an inactive dimension repeating its first child so every load is in
// bounds and unconditional) and the arithmetic and store (RestrictStore), so a lane can issue
// the loads of several cells before any store.
struct RestrictIn {
  double x[8];
  std::int64_t out;
  int c;
};

// active dimension q: every dimension when ALL3 (known at compile time), else the table's
template <bool ALL3>
__device__ __forceinline__ int Act(const ProResParams &p, int q) {
  return ALL3 ? 1 : p.act[q];
}

template <bool ALL3>
__device__ __forceinline__ RestrictIn RestrictLoad(const Region &R, const Cell &d,
                                                   const ProResParams &p) {
  const int ci = d.i + R.lo[0], cj = d.j + R.lo[1], ck = d.k + R.lo[2];
  const int a0 = Act<ALL3>(p, 0), a1 = Act<ALL3>(p, 1), a2 = Act<ALL3>(p, 2);
  const int fi = a0 ? (ci - p.clo[0]) * 2 + p.flo[0] : p.flo[0];
  const int fj = a1 ? (cj - p.clo[1]) * 2 + p.flo[1] : p.flo[1];
  const int fk = a2 ? (ck - p.clo[2]) * 2 + p.flo[2] : p.flo[2];
  const gdouble *f = Glob(R.a) + static_cast<std::int64_t>(d.c) * R.astride[2];
  RestrictIn in;
#pragma unroll
  for (int ok = 0; ok < 2; ++ok)
#pragma unroll
    for (int oj = 0; oj < 2; ++oj)
#pragma unroll
      for (int oi = 0; oi < 2; ++oi)
        in.x[(ok * 2 + oj) * 2 + oi] =
            f[static_cast<std::int64_t>(fk + ok * a2) * R.astride[1] +
              static_cast<std::int64_t>(fj + oj * a1) * R.astride[0] + (fi + oi * a0)];
  in.out = static_cast<std::int64_t>(d.c) * R.bstride[2] +
           static_cast<std::int64_t>(ck) * R.bstride[1] +
           static_cast<std::int64_t>(cj) * R.bstride[0] + ci;
  in.c = d.c;
  return in;
}

template <bool ALL3>
__device__ __forceinline__ void RestrictStore(const Region &R, const RestrictIn &in,
                                              const DeviceBlockGeom &G,
                                              const ProResParams &p) {
  const double v = G.dx[0] * G.dx[1] * G.dx[2];
  double vol[2][2][2], t[2][2][2];
#pragma unroll
  for (int ok = 0; ok < 2; ++ok)
#pragma unroll
    for (int oj = 0; oj < 2; ++oj)
#pragma unroll
      for (int oi = 0; oi < 2; ++oi) {
        const bool use = (ok == 0 || Act<ALL3>(p, 2)) && (oj == 0 || Act<ALL3>(p, 1)) &&
                         (oi == 0 || Act<ALL3>(p, 0));
        vol[ok][oj][oi] = use ? v : 0.0;
        t[ok][oj][oi] = use ? v * in.x[(ok * 2 + oj) * 2 + oi] : 0.0;
      }
  const double tvol = ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                      ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
  Glob(R.b)[in.out] = (((t[0][0][0] + t[0][1][0]) + (t[0][0][1] + t[0][1][1])) +
                       ((t[1][0][0] + t[1][1][0]) + (t[1][0][1] + t[1][1][1]))) /
                      tvol;
}

__device__ __forceinline__ double Sgn(double x) { return x < 0 ? -1.0 : 1.0; }
__device__ __forceinline__ double Min(double a, double b) { return b < a ? b : a; }

// Shared prolongation of one coarse cell into its 2x2x2 fine cells; minmod selects the
// limited slope, otherwise the one-sided slopes (ProlongateSharedLinear). ProlongLoad reads
// the cell and its two neighbours per dimension (an inactive dimension reads the cell
// itself), ProlongStore does the arithmetic and the stores.
struct ProlongIn {
  double fc, lo[3], hi[3];
  int ci[3], c;
};

template <bool ALL3>
__device__ __forceinline__ ProlongIn ProlongLoad(const Region &R, const Cell &d,
                                                 const ProResParams &p) {
  ProlongIn in;
  in.ci[0] = d.i + R.lo[0];
  in.ci[1] = d.j + R.lo[1];
  in.ci[2] = d.k + R.lo[2];
  in.c = d.c;
  const gdouble *cz = Glob(R.b) + static_cast<std::int64_t>(d.c) * R.bstride[2];
  const std::int64_t cstr[3] = {Act<ALL3>(p, 0),
                                Act<ALL3>(p, 1) * static_cast<std::int64_t>(R.bstride[0]),
                                Act<ALL3>(p, 2) * static_cast<std::int64_t>(R.bstride[1])};
  const std::int64_t cc = static_cast<std::int64_t>(in.ci[2]) * R.bstride[1] +
                          static_cast<std::int64_t>(in.ci[1]) * R.bstride[0] + in.ci[0];
  in.fc = cz[cc];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    in.lo[q] = cz[cc - cstr[q]];
    in.hi[q] = cz[cc + cstr[q]];
  }
  return in;
}

template <bool ALL3>
__device__ __forceinline__ void ProlongStore(const Region &R, const ProlongIn &in,
                                             const DeviceBlockGeom &G, const ProResParams &p,
                                             bool minmod) {
  int fi[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    fi[q] = Act<ALL3>(p, q) ? (in.ci[q] - p.clo[q]) * 2 + p.flo[q] : p.flo[q];
  gdouble *fz = Glob(R.a) + static_cast<std::int64_t>(in.c) * R.astride[2];
  const double fc = in.fc;
  double gm[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, dfm[3] = {0, 0, 0}, dfp[3] = {0, 0, 0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (!Act<ALL3>(p, q)) continue;
    const double xm = G.cx0[q] + (in.ci[q] - 1 + 0.5) * G.cdx[q];
    const double xc = G.cx0[q] + (in.ci[q] + 0.5) * G.cdx[q];
    const double xp = G.cx0[q] + (in.ci[q] + 1 + 0.5) * G.cdx[q];
    const double dxm = xc - xm, dxp = xp - xc;
    const double fxm = G.x0[q] + (fi[q] + 0.5) * G.dx[q];
    const double fxp = G.x0[q] + (fi[q] + 1 + 0.5) * G.dx[q];
    dfm[q] = xc - fxm;
    dfp[q] = fxp - xc;
    const double gxm = (fc - in.lo[q]) / dxm;
    const double gxp = (in.hi[q] - fc) / dxp;
    if (minmod) {
      const double gc = 0.5 * (Sgn(gxm) + Sgn(gxp)) * Min(fabs(gxm), fabs(gxp));
      gm[q] = gc;
      gp[q] = gc;
    } else {
      gm[q] = gxm;
      gp[q] = gxp;
    }
  }
  const std::int64_t f0 = static_cast<std::int64_t>(fi[2]) * R.astride[1] +
                          static_cast<std::int64_t>(fi[1]) * R.astride[0] + fi[0];
  const std::int64_t sj = R.astride[0], sk = R.astride[1];
  const bool X1 = Act<ALL3>(p, 0), X2 = Act<ALL3>(p, 1), X3 = Act<ALL3>(p, 2);
  fz[f0] = fc - (gm[0] * dfm[0] + gm[1] * dfm[1] + gm[2] * dfm[2]);
  if (X1) fz[f0 + 1] = fc + (gp[0] * dfp[0] - gm[1] * dfm[1] - gm[2] * dfm[2]);
  if (X2) fz[f0 + sj] = fc - (gm[0] * dfm[0] - gp[1] * dfp[1] + gm[2] * dfm[2]);
  if (X2 && X1) fz[f0 + sj + 1] = fc + (gp[0] * dfp[0] + gp[1] * dfp[1] - gm[2] * dfm[2]);
  if (X3) fz[f0 + sk] = fc - (gm[0] * dfm[0] + gm[1] * dfm[1] - gp[2] * dfp[2]);
  if (X3 && X1) fz[f0 + sk + 1] = fc + (gp[0] * dfp[0] - gm[1] * dfm[1] + gp[2] * dfp[2]);
  if (X3 && X2) fz[f0 + sk + sj] = fc - (gm[0] * dfm[0] - gp[1] * dfp[1] - gp[2] * dfp[2]);
  if (X3 && X2 && X1)
    fz[f0 + sk + sj + 1] = fc + (gp[0] * dfp[0] + gp[1] * dfp[1] + gp[2] * dfp[2]);
}

// OP 0: restriction, 1: minmod prolongation, 2: linear prolongation; ALL3: every dimension
// active (3-D meshes). A lane takes kProRes cells of the item at a time: their loads first,
// then their arithmetic and stores (lanes past the item's end repeat its last cell, which
// writes the same value to the same place).
constexpr int kProRes = kProResItemElems / kWG;
template <int OP, bool ALL3>
__global__ void __launch_bounds__(kWG)
    ProResKernel(const Region *__restrict__ reg, const Item *__restrict__ items, int nitems,
                 const DeviceBlockGeom *__restrict__ geom, ProResParams p) {
  const Item I = items[blockIdx.x];
  const Region &R = reg[I.region];
  const DeviceBlockGeom &G = geom[R.gid];
  for (int u0 = 0; u0 * kWG < I.count; u0 += kProRes) {
    if (OP == 0) {
      RestrictIn in[kProRes];
#pragma unroll
      for (int u = 0; u < kProRes; ++u)
        in[u] = RestrictLoad<ALL3>(
            R, Decode(R, static_cast<std::uint32_t>(I.begin + Lane(I, u0 + u))), p);
#pragma unroll
      for (int u = 0; u < kProRes; ++u) RestrictStore<ALL3>(R, in[u], G, p);
    } else {
      ProlongIn in[kProRes];
#pragma unroll
      for (int u = 0; u < kProRes; ++u)
        in[u] = ProlongLoad<ALL3>(
            R, Decode(R, static_cast<std::uint32_t>(I.begin + Lane(I, u0 + u))), p);
#pragma unroll
      for (int u = 0; u < kProRes; ++u) ProlongStore<ALL3>(R, in[u], G, p, OP == 1);
    }
  }
}

int GridFor(int nitems) { return nitems; }  // one workgroup per work item of the table

// Physical boundary conditions: one workgroup per box of ghost cells. Every cell copies its
// source cell (negated for a reflected vector component). Sources lie outside every box of
// the launch, so the order of the workgroups does not matter.
__global__ void __launch_bounds__(kWG) BoundaryConditionKernel(const BcBox *__restrict__ boxes) {
  const BcBox &B = boxes[blockIdx.x];
  const int n0 = B.n[0], n01 = B.n[0] * B.n[1];
  const int cells = n01 * B.n[2];
  const std::int64_t cs = static_cast<std::int64_t>(B.e[0]) * B.e[1] * B.e[2];
  for (int t = threadIdx.x; t < cells * B.ncomp; t += kWG) {
    const int c = t / cells, r = t % cells;
    const int i = B.lo[0] + r % n0;
    const int j = B.lo[1] + (r % n01) / n0;
    const int k = B.lo[2] + r / n01;
    const int si = B.a[0] + B.b[0] * i, sj = B.a[1] + B.b[1] * j, sk = B.a[2] + B.b[2] * k;
    double *q = B.slab + c * cs;
    double v = q[(static_cast<std::int64_t>(sk) * B.e[1] + sj) * B.e[0] + si];
    // the faces' operations in the reference's face order (x, then y, then z)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if ((B.fixed >> d) & 1)
        v = 2.0 * 0.0 - v;
      else if (B.vec && c == d && ((B.flip >> d) & 1))
        v = -1.0 * v;
    }
    q[(static_cast<std::int64_t>(k) * B.e[1] + j) * B.e[0] + i] = v;
  }
}

}  // namespace

void BcTable::Free() {
  if (dev) (void)hipFree(dev);
  dev = nullptr;
}

pth_status BcTable::Upload() {
  Free();
  if (host.empty()) return PTH_COMPLETE;
  PTH_HIP(hipMalloc(&dev, sizeof(BcBox) * host.size()));
  PTH_HIP(hipMemcpy(dev, host.data(), sizeof(BcBox) * host.size(), hipMemcpyHostToDevice));
  return PTH_COMPLETE;
}

pth_status LaunchBoundaryConditions(const BcTable &t, hipStream_t s) {
  if (t.host.empty()) return PTH_COMPLETE;
  ProfScope prof("boundary_conditions", s);
  hipLaunchKernelGGL(BoundaryConditionKernel, dim3(static_cast<unsigned>(t.host.size())),
                     dim3(kWG), 0, s, t.dev);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

std::uint64_t Magic(int d) {
  const std::uint64_t one = 1ull << 40;
  return (one + static_cast<std::uint64_t>(d) - 1) / static_cast<std::uint64_t>(d);
}

void Table::Free() {
  if (dev) (void)hipFree(dev);
  if (dev_items) (void)hipFree(dev_items);
  dev = nullptr;
  dev_items = nullptr;
}

// Assigns region start offsets, cuts regions into work items and copies both to the device.
pth_status Table::Upload() {
  Free();
  total = 0;
  items.clear();
  // every region is cut into items of at most item_elems elements; one workgroup per item
  for (size_t q = 0; q < host.size(); ++q) {
    Region &r = host[q];
    r.start = total;
    const std::int64_t count = static_cast<std::int64_t>(r.n[0]) * r.n[1] * r.n[2] * r.nc;
    if (count >= (1ll << 24)) {
      SetError("boundary region too large for the 32-bit element decode");
      return PTH_FAIL;
    }
    const int step = std::min(std::max(item_elems, 1), kItemElems);
    for (std::int64_t b = 0; b < count; b += step)
      items.push_back(Item{static_cast<int>(q), 1, static_cast<int>(b),
                           static_cast<int>(std::min<std::int64_t>(count - b, step))});
    total += count;
  }
  if (host.empty()) return PTH_COMPLETE;
  PTH_HIP(hipMalloc(&dev, sizeof(Region) * host.size()));
  PTH_HIP(hipMalloc(&dev_items, sizeof(Item) * items.size()));
  PTH_HIP(hipMemcpy(dev, host.data(), sizeof(Region) * host.size(), hipMemcpyHostToDevice));
  PTH_HIP(hipMemcpy(dev_items, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice));
  return PTH_COMPLETE;
}

pth_status LaunchPack(const Table &t, double *, hipStream_t s, bool to_buffer, int mode) {
  if (t.total == 0) return PTH_COMPLETE;
  ProfScope prof(to_buffer ? "pack" : "unpack", s);
  if (mode == PTH_PACK_PER_BUFFER) {
    for (const auto &R : t.host) {
      const std::int64_t count = static_cast<std::int64_t>(R.n[0]) * R.n[1] * R.n[2] * R.nc;
      hipLaunchKernelGGL(PackOneKernel, dim3(static_cast<unsigned>((count + kWG - 1) / kWG)),
                         dim3(kWG), 0, s, R, to_buffer ? 1 : 0);
    }
  } else {
    const int n = static_cast<int>(t.items.size());
    if (to_buffer)
      hipLaunchKernelGGL(PackKernel<true>, dim3(GridFor(n)), dim3(kWG), 0, s, t.dev, t.dev_items);
    else
      hipLaunchKernelGGL(PackKernel<false>, dim3(GridFor(n)), dim3(kWG), 0, s, t.dev, t.dev_items);
  }
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status LaunchCopy(const Table &t, hipStream_t s) {
  if (t.total == 0) return PTH_COMPLETE;
  ProfScope prof("remesh_copy", s);
  const int n = static_cast<int>(t.items.size());
  hipLaunchKernelGGL(CopyKernel, dim3(GridFor(n)), dim3(kWG), 0, s, t.dev, t.dev_items);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

bool CanExchangeDirect(const Table &pack, const Table &unpack) {
  return pack.dev && unpack.dev && pack.host.size() == unpack.host.size() &&
         pack.total == unpack.total;
}

pth_status LaunchExchangeDirect(const Table &pack, const Table &unpack, hipStream_t s) {
  if (pack.total == 0) return PTH_COMPLETE;
  if (!CanExchangeDirect(pack, unpack)) {
    SetError("internal: direct exchange needs paired send and receive regions");
    return PTH_FAIL;
  }
  ProfScope prof("exchange_direct", s);
  hipLaunchKernelGGL(ExchangeRegionKernel, dim3(GridFor(static_cast<int>(pack.items.size()))),
                     dim3(kWG), 0, s, pack.dev, unpack.dev, pack.dev_items);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

bool ShellSupported(const Topology &t) {
  const Extent &x = t.fine;
  if (t.ndim != 3 || x.n[0] != x.n[1] || x.n[0] != x.n[2] || x.g[0] != x.g[1] ||
      x.g[0] != x.g[2])
    return false;
  const int n = x.n[0], g = x.g[0];
  return (n == 8 || n == 16 || n == 32) && g >= 2 && g <= 4;
}

pth_status LaunchShell(const Topology &t, const std::vector<ShellBrick> &bricks,
                       const int *same, const int *order, hipStream_t s) {
  const int nb = static_cast<int>(t.blocks.size());
  if (bricks.empty() || nb == 0) return PTH_COMPLETE;
  if (!ShellSupported(t) || !same || !order) {
    SetError("internal: ghost-shell exchange needs cubic 8/16/32 blocks and 2-4 ghosts");
    return PTH_FAIL;
  }
  ProfScope prof("exchange_shell", s);
  const int n = t.fine.n[0], g = t.fine.g[0];
  for (size_t q0 = 0; q0 < bricks.size(); q0 += kShellBricks) {
    ShellArgs A{};
    const int nq = static_cast<int>(std::min<size_t>(kShellBricks, bricks.size() - q0));
    for (int q = 0; q < nq; ++q) {
      A.brick[q] = bricks[q0 + q].base;
      A.stride[q] = bricks[q0 + q].stride;
    }
    A.same = same;
    A.order = order;
    A.nblocks = nb;
    const dim3 grid(static_cast<unsigned>(nq) * nb);
#define PTH_SHELL(NN, GG)                                                              \
  if (n == NN && g == GG) hipLaunchKernelGGL((ShellKernel<NN, GG>), grid, dim3(kWG), 0, s, A);
    PTH_SHELL(8, 2) PTH_SHELL(8, 3) PTH_SHELL(8, 4)
    PTH_SHELL(16, 2) PTH_SHELL(16, 3) PTH_SHELL(16, 4)
    PTH_SHELL(32, 2) PTH_SHELL(32, 3) PTH_SHELL(32, 4)
#undef PTH_SHELL
    PTH_HIP(hipGetLastError());
  }
  return PTH_COMPLETE;
}

static ProResParams Params(const Topology &topo) {
  ProResParams p;
  for (int d = 0; d < 3; ++d) {
    p.flo[d] = topo.fine.lo(d);
    p.clo[d] = topo.coarse.lo(d);
    p.act[d] = topo.active[d] ? 1 : 0;
  }
  return p;
}

template <int OP>
static void LaunchProRes(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                         hipStream_t s) {
  const int n = static_cast<int>(t.items.size());
  const ProResParams P = Params(topo);
  if (P.act[0] && P.act[1] && P.act[2])
    hipLaunchKernelGGL((ProResKernel<OP, true>), dim3(GridFor(n)), dim3(kWG), 0, s, t.dev,
                       t.dev_items, n, geom, P);
  else
    hipLaunchKernelGGL((ProResKernel<OP, false>), dim3(GridFor(n)), dim3(kWG), 0, s, t.dev,
                       t.dev_items, n, geom, P);
}

pth_status LaunchRestrictT(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                           hipStream_t s) {
  if (t.total == 0) return PTH_COMPLETE;
  ProfScope prof("restrict", s);
  LaunchProRes<0>(t, topo, geom, s);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status LaunchProlongT(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                          hipStream_t s, bool minmod) {
  if (t.total == 0) return PTH_COMPLETE;
  ProfScope prof("prolong", s);
  if (minmod)
    LaunchProRes<1>(t, topo, geom, s);
  else
    LaunchProRes<2>(t, topo, geom, s);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

}  // namespace pth
