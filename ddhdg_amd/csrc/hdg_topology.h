// hdg_topology.h -- structured box mesh numbering and the trace-matrix block
// pattern (host only).
//
// Cells: c = ix + nx*(iy + ny*iz).  Faces: axis-major (all faces of normal x,
// then y, then z), lexicographic on each face grid.  Local face lf = 2a+s of a
// cell is its face of normal a on side s, whose boundary id is lf when it
// lies on the domain boundary (GridGenerator::subdivided_hyper_rectangle with
// colorize=true).  Two faces couple in the trace matrix when they belong to
// one cell (DoFTools::make_sparsity_pattern on the trace space,
// reference src/np_solver.cpp:644-652).
#pragma once
#include <vector>

namespace hdg
{
  struct Topology
  {
    int       d    = 2;
    int       n[3] = {1, 1, 1};
    long long ncells = 0, nfaces = 0;
    long long face_off[4]     = {0, 0, 0, 0};
    long long face_dims[3][3] = {};
    // block pattern: the faces coupled with face f are
    // blk_cols[blkptr[f] .. blkptr[f+1]), sorted; nbk[f] of them
    std::vector<long long>     blkptr, blk_cols;
    std::vector<int>           nbk;
    // cslot[(c*nfc + a)*nfc + b]: position of cell c's face b among the
    // faces coupled with its face a
    std::vector<unsigned char> cslot;

    int       nfc() const { return 2 * d; }
    long long face_index(int a, const int pos[3]) const
    {
      const long long *m = face_dims[a];
      return face_off[a] + pos[0] + m[0] * (pos[1] + m[1] * (long long)pos[2]);
    }
    void cell_coords(long long c, int ic[3]) const
    {
      ic[0]       = (int)(c % n[0]);
      long long r = c / n[0];
      ic[1]       = (int)(r % n[1]);
      ic[2]       = (int)(r / n[1]);
    }
    long long cell_face(long long c, int lf) const
    {
      int ic[3];
      cell_coords(c, ic);
      ic[lf / 2] += lf % 2;
      return face_index(lf / 2, ic);
    }
    int boundary_id(long long c, int lf) const
    {
      int ic[3];
      cell_coords(c, ic);
      const int a = lf / 2, s = lf % 2;
      if ((s == 0 && ic[a] == 0) || (s == 1 && ic[a] == n[a] - 1))
        return lf;
      return -1;
    }
  };

  // numbering and block pattern of a d-dimensional n[0] x n[1] (x n[2]) grid
  void build_topology(int d, const int n[3], Topology &T);
} // namespace hdg
