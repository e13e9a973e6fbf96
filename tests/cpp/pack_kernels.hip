// Package kernels written the way the reference writes them, against a pack instead of raw
// pointers: burgers_package::CalculateDerived (benchmarks/burgers/burgers_package.cpp:143-167,
// PackVariables({"derived", "U"})) and Update::FluxDivergence (src/interface/update.cpp:63-84
// with FluxDivHelper, update.hpp:43-57; PackVariablesAndFluxes). Built with
// -ffp-contract=off, so each is reproducible bit for bit by an IEEE CPU evaluation. The GPU
// tests load this library with ctypes and compare against the oracle and against the
// library's own kernels.
#include <hip/hip_runtime.h>

#include "parthenon_amd_pack.hpp"

using parthenon::SparsePack;

namespace {

// one workgroup row per (b, k, j), lanes along i: the par_for_outer / par_for_inner shape
__global__ void CalculateDerivedKernel(SparsePack v) {
  const int nj = v.je() - v.js() + 1, nk = v.ke() - v.ks() + 1;
  const int row = blockIdx.x;
  const int b = row / (nj * nk), k = v.ks() + (row / nj) % nk, j = v.js() + row % nj;
  for (int i = v.is() + threadIdx.x; i <= v.ie(); i += blockDim.x) {
    const double u1 = v(b, 1, k, j, i), u2 = v(b, 2, k, j, i), u3 = v(b, 3, k, j, i);
    const double d = v(b, 4, k, j, i);
    v(b, 0, k, j, i) = 0.5 * d * (u1 * u1 + u2 * u2 + u3 * u3);
  }
}

__global__ void FluxDivergenceKernel(SparsePack vin, SparsePack dudt, int ndim) {
  const int nj = vin.je() - vin.js() + 1, nk = vin.ke() - vin.ks() + 1;
  const int nv = vin.GetMaxNumberOfVars();
  const int row = blockIdx.x;  // (b, l, k, j)
  const int j = vin.js() + row % nj;
  const int k = vin.ks() + (row / nj) % nk;
  const int l = (row / (nj * nk)) % nv;
  const int b = row / (nj * nk * nv);
  for (int i = vin.is() + threadIdx.x; i <= vin.ie(); i += blockDim.x) {
    const double a1 = vin.FaceArea(b, 1), a2 = vin.FaceArea(b, 2), a3 = vin.FaceArea(b, 3);
    double du = (a1 * vin.flux(b, 1, l, k, j, i + 1) - a1 * vin.flux(b, 1, l, k, j, i));
    if (ndim >= 2) du += (a2 * vin.flux(b, 2, l, k, j + 1, i) - a2 * vin.flux(b, 2, l, k, j, i));
    if (ndim == 3) du += (a3 * vin.flux(b, 3, l, k + 1, j, i) - a3 * vin.flux(b, 3, l, k, j, i));
    dudt(b, l, k, j, i) = -du / vin.CellVolume(b);
  }
}

// Xc of every interior cell of pack index 0, as an [b][k][j][i] interior array per direction
__global__ void CoordinatesKernel(SparsePack v, double *out) {
  const int ni = v.ie() - v.is() + 1, nj = v.je() - v.js() + 1, nk = v.ke() - v.ks() + 1;
  const long n = static_cast<long>(v.GetNBlocks()) * nk * nj * ni;
  for (long t = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<long>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(t % ni), j = static_cast<int>((t / ni) % nj);
    const int k = static_cast<int>((t / ni / nj) % nk), b = static_cast<int>(t / ni / nj / nk);
    out[t] = v.Xc(b, 1, v.is() + i);
    out[n + t] = v.Xc(b, 2, v.js() + j);
    out[2 * n + t] = v.Xc(b, 3, v.ks() + k);
  }
}

int Fail() { return PTH_FAIL; }

}  // namespace

extern "C" {

// CalculateDerived over md through the pack {"derived", "U"}
int pack_calculate_derived(pth_meshdata *md) {
  const char *vars[] = {"derived", "U"};
  pth_pack *p = pth_pack_get(md, vars, 2, 0);
  if (!p) return Fail();
  SparsePack v(p);
  int lo = -1, hi = -1;
  if (pth_pack_bounds(p, 1, &lo, &hi) != PTH_COMPLETE || lo != 1) return Fail();
  const int rows = v.GetNBlocks() * (v.ke() - v.ks() + 1) * (v.je() - v.js() + 1);
  hipLaunchKernelGGL(CalculateDerivedKernel, dim3(rows), dim3(64), 0, 0, v);
  return hipDeviceSynchronize() == hipSuccess ? PTH_COMPLETE : PTH_FAIL;
}

// FluxDivergence of every WithFluxes variable ("U") of in into dudt
int pack_flux_divergence(pth_meshdata *in, pth_meshdata *dudt, int ndim) {
  const char *vars[] = {"U"};
  pth_pack *pi = pth_pack_get(in, vars, 1, PTH_PACK_WITH_FLUXES);
  pth_pack *po = pth_pack_get(dudt, vars, 1, 0);
  if (!pi || !po) return Fail();
  SparsePack vin(pi), vout(po);
  const int rows = vin.GetNBlocks() * vin.GetMaxNumberOfVars() * (vin.ke() - vin.ks() + 1) *
                   (vin.je() - vin.js() + 1);
  hipLaunchKernelGGL(FluxDivergenceKernel, dim3(rows), dim3(64), 0, 0, vin, vout, ndim);
  return hipDeviceSynchronize() == hipSuccess ? PTH_COMPLETE : PTH_FAIL;
}

// cell-centre coordinates through the pack of field (fine or coarse bricks) into host memory
int pack_coordinates(pth_meshdata *md, const char *field, int coarse, double *host, long n) {
  const char *vars[] = {field};
  pth_pack *p = pth_pack_get(md, vars, 1, coarse ? PTH_PACK_COARSE : 0);
  if (!p) return Fail();
  SparsePack v(p);
  const long cells = static_cast<long>(v.GetNBlocks()) * (v.ie() - v.is() + 1) *
                     (v.je() - v.js() + 1) * (v.ke() - v.ks() + 1);
  if (n != 3 * cells) return Fail();
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * n) != hipSuccess) return Fail();
  hipLaunchKernelGGL(CoordinatesKernel, dim3(256), dim3(256), 0, 0, v, d);
  const bool ok = hipMemcpy(host, d, sizeof(double) * n, hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(d);
  return ok ? PTH_COMPLETE : PTH_FAIL;
}

}  // extern "C"
