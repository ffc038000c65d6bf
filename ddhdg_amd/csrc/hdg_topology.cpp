// hdg_topology.cpp -- structured box mesh numbering and trace block pattern.
#include "hdg_topology.h"

#include <algorithm>

namespace hdg
{
  void build_topology(int d, const int n[3], Topology &T)
  {
    T.d      = d;
    T.ncells = 1;
    for (int a = 0; a < 3; ++a)
      {
        T.n[a] = a < d ? n[a] : 1;
        T.ncells *= T.n[a];
      }
    T.face_off[0] = 0;
    for (int a = 0; a < 3; ++a)
      {
        for (int b = 0; b < 3; ++b)
          T.face_dims[a][b] = T.n[b] + (a == b ? 1 : 0);
        const long long cnt = a < d ? T.face_dims[a][0] * T.face_dims[a][1] * T.face_dims[a][2] : 0;
        T.face_off[a + 1] = T.face_off[a] + cnt;
      }
    T.nfaces      = T.face_off[d];
    const int nfc = T.nfc();

    std::vector<std::vector<long long>> cols(T.nfaces);
    for (long long c = 0; c < T.ncells; ++c)
      for (int a = 0; a < nfc; ++a)
        {
          const long long fa = T.cell_face(c, a);
          for (int b = 0; b < nfc; ++b)
            cols[fa].push_back(T.cell_face(c, b));
        }
    T.blkptr.assign(T.nfaces + 1, 0);
    T.nbk.assign(T.nfaces, 0);
    for (long long f = 0; f < T.nfaces; ++f)
      {
        std::sort(cols[f].begin(), cols[f].end());
        cols[f].erase(std::unique(cols[f].begin(), cols[f].end()), cols[f].end());
        T.nbk[f]         = (int)cols[f].size();
        T.blkptr[f + 1] = T.blkptr[f] + T.nbk[f];
      }
    T.blk_cols.clear();
    T.blk_cols.reserve(T.blkptr[T.nfaces]);
    for (long long f = 0; f < T.nfaces; ++f)
      T.blk_cols.insert(T.blk_cols.end(), cols[f].begin(), cols[f].end());
    T.cslot.assign(T.ncells * nfc * nfc, 0);
    for (long long c = 0; c < T.ncells; ++c)
      for (int a = 0; a < nfc; ++a)
        {
          const long long fa = T.cell_face(c, a);
          const long long *b0 = T.blk_cols.data() + T.blkptr[fa];
          const long long *b1 = T.blk_cols.data() + T.blkptr[fa + 1];
          for (int b = 0; b < nfc; ++b)
            {
              const long long fb = T.cell_face(c, b);
              T.cslot[(c * nfc + a) * nfc + b] = (unsigned char)(std::lower_bound(b0, b1, fb) - b0);
            }
        }
  }
} // namespace hdg
