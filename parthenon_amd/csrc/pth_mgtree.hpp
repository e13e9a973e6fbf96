// The multigrid solver on block-mesh hierarchies (pth_mgtree.hip), used by pth_mg_* for
// meshes that are not a uniform N^3 cube with FixedFace faces (or when an SRJ smoother is
// asked for).
#pragma once
#include <vector>

#include "pth_internal.hpp"

namespace pth {

class MgTree;
// nullptr with the error set on failure; fD: the face field D (3 components) or -1 for the
// constant D
MgTree *MgTreeCreate(Mesh *m, int fu, int frhs, int fD, const pth_mg_params &p, double D,
                     double alpha);
void MgTreeDestroy(MgTree *t);
int MgTreeLevels(const MgTree *t);
// on (the default): V-cycles after the first of a solver replay one captured hipGraph
void MgTreeSetGraph(MgTree *t, bool on);
// copies u, rhs (and the face field D, restricted down the levels) of md into the hierarchy
pth_status MgTreeSetup(MgTree *t, MeshData &md);
// V-cycles until the RMS residual is below the tolerance (at least one, at most max_iters);
// writes u's cells (ghosts included, stale: the caller refills them); history: the RMS
// residual before every V-cycle and after the last
pth_status MgTreeSolve(MgTree *t, MeshData &md, int *iters, std::vector<double> *history);

}  // namespace pth
