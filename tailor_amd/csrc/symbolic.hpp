// symbolic.hpp -- host symbolic analysis for the multifrontal LU of A - sigma B.
//
// The block graph of the operator is the tensor product of the two 1-D FD-q
// stencils (77 neighbours per node at q_y = 10, q_z = 6), so the analysis
// works on nodes (5x5 blocks) and uses the grid coordinates directly:
// geometric nested dissection with one-sided vertex separators computed from
// the (symmetrised) stencil adjacency.  This replaces the ordering and
// symbolic phase MUMPS / PETSc run inside PCLU (EigenSolver.cpp:61-67).
#pragma once

#include <cstdint>
#include <vector>

namespace tailor {

struct Symbolic {
  int ny = 0, nz = 0, nodes = 0;
  std::vector<int> pos;      // node -> elimination position
  std::vector<int> order;    // elimination position -> node
  // supernodes (fronts) in postorder: children before parents
  int nsn = 0;
  std::vector<int> sn_start;              // position range [sn_start[s], sn_start[s+1])
  std::vector<int> parent;                // -1 at roots
  std::vector<std::vector<int>> children;
  std::vector<int> level;                 // 0 at leaves; parent > every child
  std::vector<std::vector<int>> strct;    // update nodes of each front, ascending position
  std::vector<std::vector<int>> rel;      // per front: index of each update node in the parent's front
  std::vector<std::vector<int>> levels;   // fronts of each level
  // dof sizes of each front: pivots p, update rows c, order m = p + c
  std::vector<int> p, c, m;
  int64_t flops = 0;         // real flops of the partial factorizations (LU part)
  int64_t factor_entries = 0;

  // node-level local index of `node` in front f (-1 if absent)
  int local_index(int f, int node) const;
};

// Build the analysis.  xoff (ny x ssx) and eoff (nz x sse) are the stencil
// offsets of the bounded xi rule and the periodic eta rule.
Symbolic analyse(int ny, int nz, int ssx, int sse, const std::vector<int> &xoff,
                 const std::vector<int> &eoff, int leaf_nodes = 64);

}  // namespace tailor
