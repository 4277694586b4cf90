// batch_merge.hpp - merging independent multigrid hierarchies into one block-diagonal
// hierarchy (host side of the batched path).
//
// Problems are aligned at their finest level: merged level S-1 holds every problem's
// finest level, merged level S-1-k every problem's level n_levels-1-k, S = max n_levels.
// Within a merged level the numbering keeps the reference's conventions, so the single
// problem kernels run on it unchanged:
//   velocity: the active dofs of problem 0, 1, ... first, then their Dirichlet dofs;
//   pressure: problem by problem (P1disc is cell local, no Dirichlet split);
//   cells:    problem by problem.
// A problem whose coarsest level sits on merged level l > 0 is solved directly there; its
// cells are left out of that level's color lists (they come after the swept cells in
// cell_reorder) and have no parent (parent -1), so the transfers skip them.
#pragma once
#include <array>
#include <vector>

#include "pmg.h"

namespace pmg {

struct MergedLevel {
  pmg_level_desc d{};                    // points into the vectors below
  std::vector<int> a_rowptr, a_kcol, b_rowptr, b_kcol, bt_rowptr, bt_kcol;
  std::vector<double> a[9], b[3], bt[3];
  std::vector<int> u_begin, u_global, p_begin, p_global, parent, child;
  std::vector<double> qq_u, qq_p;
  int n_colors = 0;
  std::vector<int> ptr_colors, cell_reorder;  // swept cells by color, then the others
  std::vector<int> coarse;               // problems whose coarsest level this is
  std::vector<int> l20;                  // (offset, length) of every projected pressure block
  std::vector<int> cell_problem;         // problem of every merged cell
};

struct MergedHierarchy {
  std::vector<MergedLevel> L;            // coarsest first
  std::vector<int> clevel;               // merged level of each problem's coarsest level
  // problem dof -> merged index, on the problem's coarsest and finest levels
  std::vector<std::vector<int>> cidx, fidx;
  // [start, end) ranges of each problem in the merged finest vector (3 components x
  // {active, Dirichlet} + pressure)
  std::vector<std::vector<std::array<long, 2>>> fseg;
};

// Throws Invalid on inconsistent input (sizes, missing transfer data, unequal QQ tables).
struct Invalid {
  int code;
  const char *what;
};
void merge_hierarchies(int n_inst, const pmg_instance_desc *inst, MergedHierarchy &out);

}  // namespace pmg
