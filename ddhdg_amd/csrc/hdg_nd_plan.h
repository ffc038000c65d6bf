// hdg_nd_plan.h -- nested-dissection multifrontal plan for the trace system
// of a structured mesh (host only; the device solver is hdg_trace_solver.hip).
//
// The cell grid is bisected recursively, longest axis first, down to boxes of
// at most 2 cells per axis.  A node eliminates
//   - at a leaf: every face of its cells not shared with another box
//     (interior faces and faces on the domain boundary);
//   - above: its separator, the faces between its two children;
// and hands the Schur complement on its interface (the faces it shares with
// the rest of the mesh) to its parent.  Trace-matrix block (f, g) is
// assembled into the front of whichever of the two eliminating nodes comes
// first; both faces are in that front.
//
// Nodes of one height and one front shape form a group (rare shapes of a
// height share one group); a group's fronts are padded to one size n = S + I
// (S eliminated dofs, I interface dofs), stored column-major one after the
// other: group g's node b has its front at fbase + b*n*n and its work vector
// at wbase + b*n.  Groups are ordered by height, so every child is factored
// before its parent.  Padded eliminated dofs carry a unit diagonal; padded
// interface dofs stay zero and are not passed up.
#pragma once
#include <string>
#include <vector>

#include "hdg_topology.h"

namespace hdg
{
  struct NDGroup
  {
    int       S = 0, I = 0, n = 0, batch = 0;
    long long fbase = 0, wbase = 0, pbase = 0;
    std::vector<int>       nodes;
    // per node of the group
    std::vector<long long> pf;   // parent's front offset (-1 at the root)
    std::vector<long long> pw;   // parent's work-vector offset
    std::vector<int>       pn;   // parent's front dimension
    std::vector<int>       slot; // which child of its parent (0/1)
    std::vector<int>       pos;  // batch*I: parent front index of an interface dof (-1 pad)
    std::vector<long long> edof; // batch*S: trace dof of an eliminated index (-1 pad)
    std::vector<long long> idof; // batch*I: trace dof of an interface index (-1 pad)
  };

  struct NDPlan
  {
    int       tdf = 0;
    long long ndof = 0, nnz = 0;
    long long Ftot = 0, wtot = 0, ptot = 0;
    long long nnodes = 0;
    int       max_front = 0;
    double    flops = 0;
    std::vector<NDGroup>   groups; // by height, leaves first
    std::vector<long long> map;    // CSR entry -> front offset
    std::vector<long long> pad;    // front offsets of padded unit diagonals
    std::vector<long long> wmap;   // trace dof -> work-vector offset
  };

  // builds the plan for the trace system with tdf dofs per face (trace CSR
  // in the face-block order of the topology's block pattern); false and a
  // message on failure
  // sizes_only: stop once the groups' shapes are known (Ftot, wtot, ptot,
  // flops, max_front, nnodes): no index maps, so it is cheap at any size
  bool build_nd_plan(const Topology &topo, int tdf, NDPlan &P, std::string &err,
                     bool sizes_only = false);

  // every index against its buffer, each front entry written by at most one
  // matrix entry, every trace dof eliminated exactly once
  bool check_nd_plan(const NDPlan &P, std::string &err);

  // the plan's multifrontal arithmetic in host loops on a matrix in the
  // trace pattern (given values, or random diagonally dominant ones when
  // given is null), against dense elimination: out = {relative error of the
  // solution, tree nodes, levels}
  bool nd_plan_selftest(int d, const int n[3], int tdf, unsigned seed, const double *given,
                        double out[3], std::string &err);
} // namespace hdg
