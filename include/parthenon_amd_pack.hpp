// parthenon_amd_pack.hpp — device-side access to a pack for HIP kernels written against the
// reference's pack interface: SparsePack (src/interface/sparse_pack.hpp:53-409) and
// VariablePack / VariableFluxPack (src/interface/variable_pack.hpp), over the view that
// pth_pack_view_get fills (include/parthenon_amd.h). Include it from a HIP translation unit:
//
//   const char *vars[] = {"derived", "U"};
//   parthenon::SparsePack pack(pth_pack_get(md, vars, 2, 0));      // host
//   ...
//   __global__ void K(parthenon::SparsePack v) {                    // device
//     v(b, 0, k, j, i) = 0.5 * v(b, 4, k, j, i) * ...;
//   }
//
// The class is a plain copy of the view (trivially copyable), so it is passed to kernels by
// value as the reference's packs are captured by value in device lambdas.
#ifndef PARTHENON_AMD_PACK_HPP_
#define PARTHENON_AMD_PACK_HPP_

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "parthenon_amd.hpp"

namespace parthenon {

// Position of a requested field in a name-based pack (reference PackIdx, pack_utils.hpp):
// the field's number in the request, plus a component offset.
struct PackIdx {
  int field = 0;
  int offset = 0;
  __host__ __device__ constexpr PackIdx(int f = 0, int off = 0) : field(f), offset(off) {}
  __host__ __device__ constexpr PackIdx operator+(int c) const { return PackIdx(field, offset + c); }
};

// TopologicalElement faces of the flux accessors (reference TE::F1, F2, F3).
enum class TE : int { CC = 0, F1 = 1, F2 = 2, F3 = 3 };

class SparsePack {
 public:
  SparsePack() = default;
  explicit SparsePack(const pth_pack_view &v) : v_(v) {}
  // Descriptor.GetPack(md) of the host side: throws like PARTHENON_REQUIRE_THROWS on failure
  explicit SparsePack(const pth_pack *p) {
    if (!p || pth_pack_view_get(p, &v_) != PTH_COMPLETE)
      throw std::runtime_error(std::string("pack: ") + pth_last_error());
  }

  __host__ __device__ int GetNBlocks() const { return v_.nblocks; }
  __host__ __device__ int GetMaxNumberOfVars() const { return v_.nvar; }
  __host__ __device__ int GetNumberOfFields() const { return v_.nfields; }
  __host__ __device__ bool WithFluxes() const { return v_.flux != nullptr; }

  // GetLowerBound / GetUpperBound (sparse_pack.hpp:133-160): whole pack, or one field
  __host__ __device__ int GetLowerBound(int /*b*/) const { return 0; }
  __host__ __device__ int GetUpperBound(int /*b*/) const { return v_.nvar - 1; }
  __device__ int GetLowerBound(int /*b*/, PackIdx f) const { return v_.bounds[f.field]; }
  __device__ int GetUpperBound(int /*b*/, PackIdx f) const {
    return v_.bounds[v_.nfields + f.field];
  }
  __device__ int GetSize(int b, PackIdx f) const {
    return GetUpperBound(b, f) - GetLowerBound(b, f) + 1;
  }
  // GetIndex (sparse_pack.hpp:216-223): lower bound plus the component offset
  __device__ int GetIndex(int b, PackIdx f) const { return GetLowerBound(b, f) + f.offset; }

  // interior index range of the packed bricks (IndexShape of md->GetBoundsI/J/K)
  __host__ __device__ int is() const { return v_.is; }
  __host__ __device__ int ie() const { return v_.ie; }
  __host__ __device__ int js() const { return v_.js; }
  __host__ __device__ int je() const { return v_.je; }
  __host__ __device__ int ks() const { return v_.ks; }
  __host__ __device__ int ke() const { return v_.ke; }

  // pack(b, v, k, j, i): cell value of pack index v of block b
  __device__ double &operator()(int b, int v, int k, int j, int i) const {
    return v_.data[b * v_.nvar + v][Cell(k, j, i)];
  }
  __device__ double &operator()(int b, PackIdx f, int k, int j, int i) const {
    return (*this)(b, GetIndex(b, f), k, j, i);
  }
  __device__ double &operator()(int b, TE, int v, int k, int j, int i) const {
    return (*this)(b, v, k, j, i);
  }
  // flattened packs (PDOpt::Flatten): flat index = b * nvar + v
  __device__ double &operator()(int flat, int k, int j, int i) const {
    return v_.data[flat][Cell(k, j, i)];
  }
  // brick of (b, v), [k][j][i]
  __device__ double *Brick(int b, int v) const { return v_.data[b * v_.nvar + v]; }

  // pack.flux(b, dir, v, k, j, i): face flux of direction dir (1, 2, 3), x-face i at index i
  __device__ double &flux(int b, int dir, int v, int k, int j, int i) const {
    return v_.flux[(dir - 1) * v_.nblocks * v_.nvar + b * v_.nvar + v][Cell(k, j, i)];
  }
  __device__ double &flux(int b, int dir, PackIdx f, int k, int j, int i) const {
    return flux(b, dir, GetIndex(b, f), k, j, i);
  }
  __device__ double &flux(int b, TE te, int v, int k, int j, int i) const {
    return flux(b, static_cast<int>(te), v, k, j, i);
  }

  // Coordinates of block b (UniformCartesian, coordinates/uniform_cartesian.hpp:41-58,
  // 94-100, 166-179, 248-270)
  __device__ double Dxc(int b, int dir) const { return v_.coords[8 * b + dir - 1]; }
  __device__ double FaceArea(int b, int dir) const { return v_.coords[8 * b + 2 + dir]; }
  __device__ double CellVolume(int b) const { return v_.coords[8 * b + 6]; }
  __device__ double Xc(int b, int dir, int idx) const {
    const bool c = v_.options & PTH_PACK_COARSE;
    const double *g = v_.geom + 12 * b + (c ? 6 : 0);
    return g[dir - 1] + (idx + 0.5) * g[2 + dir];
  }

  const pth_pack_view &view() const { return v_; }

 private:
  __host__ __device__ long Cell(int k, int j, int i) const {
    return (static_cast<long>(k) * v_.ent[1] + j) * v_.ent[0] + i;
  }
  pth_pack_view v_{};
};

// MeshData::PackVariables / PackVariablesAndFluxes (mesh_data.hpp:349-430) return the same
// view; the reference's VariablePack and VariableFluxPack are index-compatible with it.
using VariablePack = SparsePack;
using VariableFluxPack = SparsePack;

}  // namespace parthenon

#endif  // PARTHENON_AMD_PACK_HPP_
