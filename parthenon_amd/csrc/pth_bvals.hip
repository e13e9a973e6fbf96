// Boundary-buffer kernels for gfx950: pack (SendBoundBufs), unpack (SetBounds), cell-average
// restriction and minmod / linear shared prolongation.
//
// Every call is ONE persistent launch that walks a device table of boundary regions of
// uneven sizes (3-wide x-face strips, 16-wide y/z strips, 27-cell corners, coarse-fine
// regions). The table flattens all regions into one element space; a workgroup takes
// consecutive 256-element chunks, so lanes of a wavefront touch consecutive buffer elements
// (fully coalesced 64-lane stores into the buffer arena) and runs of ni consecutive i in the
// field brick. A per-chunk hint bounds the binary search for the region of an element.
//
// Arithmetic follows the reference operators term by term (no FMA contraction; the library
// is built with -ffp-contract=off) so restriction is bit-exact:
//   RestrictAverage         src/prolong_restrict/pr_ops.hpp:105-164
//   ProlongateSharedGeneral src/prolong_restrict/pr_ops.hpp:166-289, GradMinMod :95-101
//   pack / unpack loops     src/bvals/comms/boundary_communication.cpp:95-139, 273-334
#include <hip/hip_runtime.h>

#include "pth_internal.hpp"

namespace pth {

namespace {

constexpr int kWG = 256;     // threads per workgroup (4 wavefronts of 64)
constexpr int kChunk = 256;  // elements per chunk of the flattened element space

struct Decoded {
  int r;  // region
  int i, j, k, c;
  std::int64_t e;  // element within the region
};

__device__ __forceinline__ int FindRegion(const std::int64_t *__restrict__ start,
                                          const int *__restrict__ hint, int nreg,
                                          std::int64_t g) {
  const std::int64_t chunk = g / kChunk;
  int lo = hint[chunk];
  int hi = hint[chunk + 1];
  if (hi >= nreg) hi = nreg - 1;
  while (lo < hi) {  // last region with start <= g
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ Decoded Decode(const Region &R, int r, std::int64_t e) {
  Decoded d;
  d.r = r;
  d.e = e;
  const std::uint64_t q0 = (static_cast<std::uint64_t>(e) * R.magic[0]) >> 40;
  d.i = static_cast<int>(e - static_cast<std::int64_t>(q0) * R.n[0]);
  const std::uint64_t q1 = (q0 * R.magic[1]) >> 40;
  d.j = static_cast<int>(q0 - q1 * R.n[1]);
  const std::uint64_t q2 = (q1 * R.magic[2]) >> 40;
  d.k = static_cast<int>(q1 - q2 * R.n[2]);
  d.c = static_cast<int>(q2);
  return d;
}

// Pack (to_buffer) or unpack one region element.
__global__ void __launch_bounds__(kWG)
    PackKernel(const Region *__restrict__ reg, const std::int64_t *__restrict__ start,
               const int *__restrict__ hint, int nreg, std::int64_t total, int to_buffer) {
  for (std::int64_t g = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; g < total;
       g += static_cast<std::int64_t>(gridDim.x) * kWG) {
    const int r = FindRegion(start, hint, nreg, g);
    const Region &R = reg[r];
    const Decoded d = Decode(R, r, g - R.start);
    const std::int64_t off = static_cast<std::int64_t>(d.c) * R.astride[2] +
                             static_cast<std::int64_t>(d.k + R.lo[2]) * R.astride[1] +
                             static_cast<std::int64_t>(d.j + R.lo[1]) * R.astride[0] +
                             (d.i + R.lo[0]);
    if (to_buffer)
      R.b[d.e] = R.a[off];
    else
      R.a[off] = R.b[d.e];
  }
}

// The plain form: one launch per region, one thread per element.
__global__ void __launch_bounds__(kWG) PackOneKernel(Region R, int to_buffer) {
  const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x;
  const std::int64_t count = static_cast<std::int64_t>(R.n[0]) * R.n[1] * R.n[2] * R.nc;
  if (e >= count) return;
  const Decoded d = Decode(R, 0, e);
  const std::int64_t off = static_cast<std::int64_t>(d.c) * R.astride[2] +
                           static_cast<std::int64_t>(d.k + R.lo[2]) * R.astride[1] +
                           static_cast<std::int64_t>(d.j + R.lo[1]) * R.astride[0] +
                           (d.i + R.lo[0]);
  if (to_buffer)
    R.b[d.e] = R.a[off];
  else
    R.a[off] = R.b[d.e];
}

struct ProResParams {
  int flo[3];  // fine interior start per dimension
  int clo[3];  // coarse interior start per dimension
  int act[3];  // active dimensions
};

// Cell-average restriction of one coarse cell from its 2x2x2 fine children; the summation
// order is the reference's fixed pairwise order.
__global__ void __launch_bounds__(kWG)
    RestrictKernel(const Region *__restrict__ reg, const std::int64_t *__restrict__ start,
                   const int *__restrict__ hint, int nreg, std::int64_t total,
                   const DeviceBlockGeom *__restrict__ geom, ProResParams p) {
  for (std::int64_t g = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; g < total;
       g += static_cast<std::int64_t>(gridDim.x) * kWG) {
    const int r = FindRegion(start, hint, nreg, g);
    const Region &R = reg[r];
    const Decoded d = Decode(R, r, g - R.start);
    const int ci = d.i + R.lo[0], cj = d.j + R.lo[1], ck = d.k + R.lo[2];
    const int fi = p.act[0] ? (ci - p.clo[0]) * 2 + p.flo[0] : p.flo[0];
    const int fj = p.act[1] ? (cj - p.clo[1]) * 2 + p.flo[1] : p.flo[1];
    const int fk = p.act[2] ? (ck - p.clo[2]) * 2 + p.flo[2] : p.flo[2];
    const DeviceBlockGeom &G = geom[R.gid];
    const double v = G.dx[0] * G.dx[1] * G.dx[2];
    const double *f = R.a + static_cast<std::int64_t>(d.c) * R.astride[2];
    double vol[2][2][2], t[2][2][2];
#pragma unroll
    for (int ok = 0; ok < 2; ++ok)
#pragma unroll
      for (int oj = 0; oj < 2; ++oj)
#pragma unroll
        for (int oi = 0; oi < 2; ++oi) {
          const bool use = (ok == 0 || p.act[2]) && (oj == 0 || p.act[1]) && (oi == 0 || p.act[0]);
          vol[ok][oj][oi] = use ? v : 0.0;
          t[ok][oj][oi] = use ? v * f[static_cast<std::int64_t>(fk + ok) * R.astride[1] +
                                      static_cast<std::int64_t>(fj + oj) * R.astride[0] +
                                      (fi + oi)]
                              : 0.0;
        }
    const double tvol = ((vol[0][0][0] + vol[0][1][0]) + (vol[0][0][1] + vol[0][1][1])) +
                        ((vol[1][0][0] + vol[1][1][0]) + (vol[1][0][1] + vol[1][1][1]));
    R.b[static_cast<std::int64_t>(d.c) * R.bstride[2] + static_cast<std::int64_t>(ck) * R.bstride[1] +
        static_cast<std::int64_t>(cj) * R.bstride[0] + ci] =
        (((t[0][0][0] + t[0][1][0]) + (t[0][0][1] + t[0][1][1])) +
         ((t[1][0][0] + t[1][1][0]) + (t[1][0][1] + t[1][1][1]))) /
        tvol;
  }
}

__device__ __forceinline__ double Sgn(double x) { return x < 0 ? -1.0 : 1.0; }
__device__ __forceinline__ double Min(double a, double b) { return b < a ? b : a; }

// Shared prolongation of one coarse cell into its 2x2x2 fine cells; minmod != 0 selects the
// limited slope, otherwise the one-sided slopes (ProlongateSharedLinear).
__global__ void __launch_bounds__(kWG)
    ProlongKernel(const Region *__restrict__ reg, const std::int64_t *__restrict__ start,
                  const int *__restrict__ hint, int nreg, std::int64_t total,
                  const DeviceBlockGeom *__restrict__ geom, ProResParams p, int minmod) {
  for (std::int64_t g = static_cast<std::int64_t>(blockIdx.x) * kWG + threadIdx.x; g < total;
       g += static_cast<std::int64_t>(gridDim.x) * kWG) {
    const int r = FindRegion(start, hint, nreg, g);
    const Region &R = reg[r];
    const Decoded d = Decode(R, r, g - R.start);
    const int ci[3] = {d.i + R.lo[0], d.j + R.lo[1], d.k + R.lo[2]};
    int fi[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) fi[q] = p.act[q] ? (ci[q] - p.clo[q]) * 2 + p.flo[q] : p.flo[q];
    const DeviceBlockGeom &G = geom[R.gid];
    const double *cz = R.b + static_cast<std::int64_t>(d.c) * R.bstride[2];
    double *fz = R.a + static_cast<std::int64_t>(d.c) * R.astride[2];
    const std::int64_t cstr[3] = {1, R.bstride[0], R.bstride[1]};
    const std::int64_t cc = static_cast<std::int64_t>(ci[2]) * R.bstride[1] +
                            static_cast<std::int64_t>(ci[1]) * R.bstride[0] + ci[0];
    const double fc = cz[cc];
    double gm[3] = {0, 0, 0}, gp[3] = {0, 0, 0}, dfm[3] = {0, 0, 0}, dfp[3] = {0, 0, 0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (!p.act[q]) continue;
      const double xm = G.cx0[q] + (ci[q] - 1 + 0.5) * G.cdx[q];
      const double xc = G.cx0[q] + (ci[q] + 0.5) * G.cdx[q];
      const double xp = G.cx0[q] + (ci[q] + 1 + 0.5) * G.cdx[q];
      const double dxm = xc - xm, dxp = xp - xc;
      const double fxm = G.x0[q] + (fi[q] + 0.5) * G.dx[q];
      const double fxp = G.x0[q] + (fi[q] + 1 + 0.5) * G.dx[q];
      dfm[q] = xc - fxm;
      dfp[q] = fxp - xc;
      const double gxm = (fc - cz[cc - cstr[q]]) / dxm;
      const double gxp = (cz[cc + cstr[q]] - fc) / dxp;
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
    const bool X1 = p.act[0], X2 = p.act[1], X3 = p.act[2];
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
}

int GridFor(std::int64_t total) {
  // persistent grid: at most 8 workgroups per CU on 256 CUs, fewer for small tables
  const std::int64_t want = (total + kWG - 1) / kWG;
  return static_cast<int>(want < 2048 ? want : 2048);
}

}  // namespace

std::uint64_t Magic(int d) {
  const std::uint64_t one = 1ull << 40;
  return (one + static_cast<std::uint64_t>(d) - 1) / static_cast<std::uint64_t>(d);
}

void Table::Free() {
  if (dev) (void)hipFree(dev);
  if (dev_start) (void)hipFree(dev_start);
  dev = nullptr;
  dev_start = nullptr;
}

// Copies regions, start offsets and per-chunk hints (stored after the starts) to the device.
pth_status Table::Upload() {
  Free();
  total = 0;
  for (auto &r : host) {
    r.start = total;
    total += static_cast<std::int64_t>(r.n[0]) * r.n[1] * r.n[2] * r.nc;
  }
  if (host.empty()) return PTH_COMPLETE;
  const std::int64_t nchunk = (total + kChunk - 1) / kChunk;
  std::vector<std::int64_t> starts(host.size());
  for (size_t q = 0; q < host.size(); ++q) starts[q] = host[q].start;
  std::vector<int> hint(nchunk + 1);
  size_t r = 0;
  for (std::int64_t c = 0; c <= nchunk; ++c) {
    const std::int64_t g = std::min<std::int64_t>(c * kChunk, total - 1);
    while (r + 1 < host.size() && starts[r + 1] <= g) ++r;
    hint[c] = static_cast<int>(r);
  }
  hint[nchunk] = static_cast<int>(host.size() - 1);
  PTH_HIP(hipMalloc(&dev, sizeof(Region) * host.size()));
  const size_t sbytes = sizeof(std::int64_t) * host.size() + sizeof(int) * hint.size();
  PTH_HIP(hipMalloc(&dev_start, sbytes));
  PTH_HIP(hipMemcpy(dev, host.data(), sizeof(Region) * host.size(), hipMemcpyHostToDevice));
  PTH_HIP(hipMemcpy(dev_start, starts.data(), sizeof(std::int64_t) * host.size(),
                    hipMemcpyHostToDevice));
  PTH_HIP(hipMemcpy(reinterpret_cast<char *>(dev_start) + sizeof(std::int64_t) * host.size(),
                    hint.data(), sizeof(int) * hint.size(), hipMemcpyHostToDevice));
  return PTH_COMPLETE;
}

static const int *Hints(const Table &t) {
  return reinterpret_cast<const int *>(reinterpret_cast<const char *>(t.dev_start) +
                                       sizeof(std::int64_t) * t.host.size());
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
    hipLaunchKernelGGL(PackKernel, dim3(GridFor(t.total)), dim3(kWG), 0, s, t.dev, t.dev_start,
                       Hints(t), static_cast<int>(t.host.size()), t.total, to_buffer ? 1 : 0);
  }
  PTH_HIP(hipGetLastError());
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

pth_status LaunchRestrictT(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                           hipStream_t s) {
  if (t.total == 0) return PTH_COMPLETE;
  ProfScope prof("restrict", s);
  hipLaunchKernelGGL(RestrictKernel, dim3(GridFor(t.total)), dim3(kWG), 0, s, t.dev, t.dev_start,
                     Hints(t), static_cast<int>(t.host.size()), t.total, geom, Params(topo));
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

pth_status LaunchProlongT(const Table &t, const Topology &topo, const DeviceBlockGeom *geom,
                          hipStream_t s, bool minmod) {
  if (t.total == 0) return PTH_COMPLETE;
  ProfScope prof("prolong", s);
  hipLaunchKernelGGL(ProlongKernel, dim3(GridFor(t.total)), dim3(kWG), 0, s, t.dev, t.dev_start,
                     Hints(t), static_cast<int>(t.host.size()), t.total, geom, Params(topo),
                     minmod ? 1 : 0);
  PTH_HIP(hipGetLastError());
  return PTH_COMPLETE;
}

}  // namespace pth
