// hdg_nd_plan.cpp -- nested-dissection multifrontal plan (host only).
#include "hdg_nd_plan.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <unordered_map>

namespace hdg
{
  namespace
  {
    struct Box
    {
      int lo[3], hi[3];
    };

    struct Node
    {
      Box                    box;
      int                    parent = -1, slot = 0, height = 0, split = -1;
      std::vector<long long> elim, intf; // faces
      int                    bidx = -1;
    };

    int longest_axis(const Box &b, int d)
    {
      int ax = 0;
      for (int a = 1; a < d; ++a)
        if (b.hi[a] - b.lo[a] > b.hi[ax] - b.lo[ax])
          ax = a;
      return ax;
    }

    int build_tree(std::vector<Node> &nodes, int d, const Box &b, int parent, int slot)
    {
      const int id = (int)nodes.size();
      nodes.emplace_back();
      nodes[id].box    = b;
      nodes[id].parent = parent;
      nodes[id].slot   = slot;
      const int ax     = longest_axis(b, d);
      if (b.hi[ax] - b.lo[ax] <= 2)
        return id; // leaf
      const int m  = b.lo[ax] + (b.hi[ax] - b.lo[ax]) / 2;
      Box       b0 = b, b1 = b;
      b0.hi[ax]    = m;
      b1.lo[ax]    = m;
      nodes[id].split = m;
      const int c0    = build_tree(nodes, d, b0, id, 0);
      const int c1    = build_tree(nodes, d, b1, id, 1);
      nodes[id].height = 1 + std::max(nodes[c0].height, nodes[c1].height);
      return id;
    }

    // faces of normal a on plane pa inside box b
    template <class F>
    void plane_faces(const Topology &T, const Box &b, int a, int pa, F &&fn)
    {
      int       pos[3];
      const int o1 = (a + 1) % 3, o2 = (a + 2) % 3;
      pos[a]       = pa;
      for (pos[o2] = b.lo[o2]; pos[o2] < b.hi[o2]; ++pos[o2])
        for (pos[o1] = b.lo[o1]; pos[o1] < b.hi[o1]; ++pos[o1])
          fn(T.face_index(a, pos));
    }
  } // namespace

  bool build_nd_plan(const Topology &T, int tdf, NDPlan &P, std::string &err, bool sizes_only)
  {
    P        = NDPlan();
    P.tdf    = tdf;
    P.ndof   = T.nfaces * tdf;
    P.nnz    = T.blkptr[T.nfaces] * (long long)tdf * tdf;
    const int d = T.d;

    std::vector<Node> nodes;
    Box               root;
    for (int a = 0; a < 3; ++a)
      {
        root.lo[a] = 0;
        root.hi[a] = T.n[a];
      }
    build_tree(nodes, d, root, -1, 0);
    P.nnodes = (long long)nodes.size();

    std::vector<int> elim_node(T.nfaces, -1), elim_pos(T.nfaces, -1);
    for (int id = 0; id < (int)nodes.size(); ++id)
      {
        Node      &N = nodes[id];
        const Box &b = N.box;
        if (N.split < 0)
          {
            for (int a = 0; a < d; ++a)
              for (int pa = b.lo[a]; pa <= b.hi[a]; ++pa)
                {
                  const bool shared = (pa == b.lo[a] || pa == b.hi[a]) && pa > 0 && pa < T.n[a];
                  if (!shared)
                    plane_faces(T, b, a, pa, [&](long long f) { N.elim.push_back(f); });
                }
          }
        else
          plane_faces(T, b, longest_axis(b, d), N.split,
                      [&](long long f) { N.elim.push_back(f); });
        for (int a = 0; a < d; ++a)
          for (int pa : {b.lo[a], b.hi[a]})
            if (pa > 0 && pa < T.n[a])
              plane_faces(T, b, a, pa, [&](long long f) { N.intf.push_back(f); });
        for (size_t e = 0; e < N.elim.size(); ++e)
          {
            const long long f = N.elim[e];
            if (elim_node[f] >= 0)
              {
                err = "nested dissection: face eliminated twice";
                return false;
              }
            elim_node[f] = id;
            elim_pos[f]  = (int)e;
          }
      }
    for (long long f = 0; f < T.nfaces; ++f)
      if (elim_node[f] < 0)
        {
          err = "nested dissection: face never eliminated";
          return false;
        }

    // groups: nodes of one height with one front shape (eliminated and
    // interface faces); shapes shared by fewer than kMinClass nodes of a
    // height go to one padded group of that height
    constexpr int kMinClass = 32;
    int hmax = 0;
    for (const Node &N : nodes)
      hmax = std::max(hmax, N.height);
    std::vector<int> group_of_node(nodes.size(), -1);
    for (int h = 0; h <= hmax; ++h)
      {
        std::vector<std::pair<std::pair<int, int>, int>> shape; // ((S, I), node)
        for (int id = 0; id < (int)nodes.size(); ++id)
          if (nodes[id].height == h)
            shape.push_back({{(int)nodes[id].elim.size(), (int)nodes[id].intf.size()}, id});
        std::sort(shape.begin(), shape.end());
        int rest = -1;
        for (size_t a = 0; a < shape.size();)
          {
            size_t e = a;
            while (e < shape.size() && shape[e].first == shape[a].first)
              ++e;
            int g;
            if ((int)(e - a) >= kMinClass || (a == 0 && e == shape.size()))
              {
                g = (int)P.groups.size();
                P.groups.emplace_back();
              }
            else
              {
                if (rest < 0)
                  {
                    rest = (int)P.groups.size();
                    P.groups.emplace_back();
                  }
                g = rest;
              }
            for (size_t k = a; k < e; ++k)
              group_of_node[shape[k].second] = g;
            a = e;
          }
      }
    for (int id = 0; id < (int)nodes.size(); ++id)
      {
        Node    &N = nodes[id];
        NDGroup &G = P.groups[group_of_node[id]];
        N.bidx     = (int)G.nodes.size();
        G.nodes.push_back(id);
        G.S = std::max(G.S, (int)N.elim.size() * tdf);
        G.I = std::max(G.I, (int)N.intf.size() * tdf);
      }
    for (NDGroup &G : P.groups)
      {
        G.n     = G.S + G.I;
        G.batch = (int)G.nodes.size();
        G.fbase = P.Ftot;
        G.wbase = P.wtot;
        G.pbase = P.ptot;
        P.Ftot += (long long)G.batch * G.n * G.n;
        P.wtot += (long long)G.batch * G.n;
        P.ptot += (long long)G.batch * G.S;
        const double S = G.S, I = G.I;
        P.flops += G.batch * (2.0 / 3.0 * S * S * S + 2.0 * S * S * I + 2.0 * S * I * I);
        P.max_front = std::max(P.max_front, G.n);
      }
    if ((long long)P.max_front * P.max_front > (long long)INT32_MAX)
      {
        err = "nested dissection: front too large";
        return false;
      }
    if (sizes_only)
      return true;

    std::vector<std::unordered_map<long long, int>> intf_pos(nodes.size());
    for (int id = 0; id < (int)nodes.size(); ++id)
      for (size_t i = 0; i < nodes[id].intf.size(); ++i)
        intf_pos[id][nodes[id].intf[i]] = (int)i;
    auto group_of = [&](int id) -> const NDGroup & { return P.groups[group_of_node[id]]; };
    auto front_index = [&](int id, long long f) -> long long {
      if (elim_node[f] == id)
        return (long long)elim_pos[f] * tdf;
      const auto it = intf_pos[id].find(f);
      if (it == intf_pos[id].end())
        return -1;
      return group_of(id).S + (long long)it->second * tdf;
    };
    auto front_base = [&](int id) -> long long {
      const NDGroup &G = group_of(id);
      return G.fbase + (long long)nodes[id].bidx * G.n * G.n;
    };

    // matrix entries -> fronts
    P.map.assign(P.nnz, -1);
    for (long long f = 0; f < T.nfaces; ++f)
      {
        const long long nb = T.nbk[f];
        for (long long bb = T.blkptr[f]; bb < T.blkptr[f + 1]; ++bb)
          {
            const long long g  = T.blk_cols[bb];
            const int       Nf = elim_node[f], Ng = elim_node[g];
            const int       N  = nodes[Nf].height <= nodes[Ng].height ? Nf : Ng;
            const long long ri = front_index(N, f), ci = front_index(N, g);
            if (ri < 0 || ci < 0)
              {
                err = "nested dissection: matrix block outside its front";
                return false;
              }
            const long long nN = group_of(N).n, fb = front_base(N);
            for (int r = 0; r < tdf; ++r)
              for (int c = 0; c < tdf; ++c)
                P.map[T.blkptr[f] * tdf * tdf + (long long)r * nb * tdf +
                      (bb - T.blkptr[f]) * tdf + c] = fb + (ri + r) + (ci + c) * nN;
          }
      }
    P.wmap.assign(P.ndof, -1);
    for (int id = 0; id < (int)nodes.size(); ++id)
      {
        const Node    &N = nodes[id];
        const NDGroup &G = group_of(id);
        for (int j = (int)N.elim.size() * tdf; j < G.S; ++j)
          P.pad.push_back(front_base(id) + j + (long long)j * G.n);
        for (size_t e = 0; e < N.elim.size(); ++e)
          for (int r = 0; r < tdf; ++r)
            P.wmap[N.elim[e] * tdf + r] =
              G.wbase + (long long)N.bidx * G.n + (long long)e * tdf + r;
      }

    // per group: links to the parents, dof lists
    for (NDGroup &G : P.groups)
      {
        G.pf.assign(G.batch, -1);
        G.pw.assign(G.batch, 0);
        G.pn.assign(G.batch, 0);
        G.slot.assign(G.batch, 0);
        G.pos.assign((size_t)G.batch * G.I, -1);
        G.edof.assign((size_t)G.batch * G.S, -1);
        G.idof.assign((size_t)G.batch * G.I, -1);
        for (int b = 0; b < G.batch; ++b)
          {
            const int   id = G.nodes[b];
            const Node &N  = nodes[id];
            G.slot[b]      = N.slot;
            for (size_t e = 0; e < N.elim.size(); ++e)
              for (int r = 0; r < tdf; ++r)
                G.edof[(size_t)b * G.S + e * tdf + r] = N.elim[e] * tdf + r;
            for (size_t i = 0; i < N.intf.size(); ++i)
              for (int r = 0; r < tdf; ++r)
                G.idof[(size_t)b * G.I + i * tdf + r] = N.intf[i] * tdf + r;
            if (N.parent < 0)
              continue;
            const NDGroup &PG = group_of(N.parent);
            G.pf[b]           = front_base(N.parent);
            G.pw[b]           = PG.wbase + (long long)nodes[N.parent].bidx * PG.n;
            G.pn[b]           = PG.n;
            for (size_t i = 0; i < N.intf.size(); ++i)
              {
                const long long fi = front_index(N.parent, N.intf[i]);
                if (fi < 0)
                  {
                    err = "nested dissection: interface face outside the parent front";
                    return false;
                  }
                for (int r = 0; r < tdf; ++r)
                  G.pos[(size_t)b * G.I + i * tdf + r] = (int)(fi + r);
              }
          }
      }
    return check_nd_plan(P, err);
  }

  bool check_nd_plan(const NDPlan &P, std::string &err)
  {
    std::vector<unsigned char> used(P.Ftot, 0);
    for (long long v : P.map)
      {
        if (v < 0 || v >= P.Ftot)
          {
            err = "plan check: matrix entry maps outside the fronts";
            return false;
          }
        if (used[v]++)
          {
            err = "plan check: two matrix entries share a front entry";
            return false;
          }
      }
    for (long long v : P.pad)
      if (v < 0 || v >= P.Ftot || used[v]++)
        {
          err = "plan check: padded diagonal outside the fronts or on a matrix entry";
          return false;
        }
    std::vector<unsigned char> seen(P.ndof, 0);
    for (long long v : P.wmap)
      if (v < 0 || v >= P.wtot)
        {
          err = "plan check: work-vector index out of range";
          return false;
        }
    for (const NDGroup &G : P.groups)
      {
        if (G.fbase + (long long)G.batch * G.n * G.n > P.Ftot ||
            G.wbase + (long long)G.batch * G.n > P.wtot ||
            G.pbase + (long long)G.batch * G.S > P.ptot)
          {
            err = "plan check: group outside the buffers";
            return false;
          }
        for (int b = 0; b < G.batch; ++b)
          {
            if (G.pf[b] >= 0 && (G.pf[b] + (long long)G.pn[b] * G.pn[b] > P.Ftot ||
                                 G.pw[b] + G.pn[b] > P.wtot))
              {
                err = "plan check: parent outside the buffers";
                return false;
              }
            for (int r = 0; r < G.I; ++r)
              {
                const int v = G.pos[(size_t)b * G.I + r];
                if (v >= 0 && (G.pf[b] < 0 || v >= G.pn[b]))
                  {
                    err = "plan check: extend-add index outside the parent front";
                    return false;
                  }
                const long long dof = G.idof[(size_t)b * G.I + r];
                if (dof >= P.ndof)
                  {
                    err = "plan check: interface dof out of range";
                    return false;
                  }
              }
            for (int j = 0; j < G.S; ++j)
              {
                const long long dof = G.edof[(size_t)b * G.S + j];
                if (dof >= P.ndof)
                  {
                    err = "plan check: eliminated dof out of range";
                    return false;
                  }
                if (dof >= 0 && seen[dof]++)
                  {
                    err = "plan check: dof eliminated twice";
                    return false;
                  }
              }
          }
      }
    for (long long i = 0; i < P.ndof; ++i)
      if (!seen[i])
        {
          err = "plan check: dof never eliminated";
          return false;
        }
    return true;
  }
} // namespace hdg

namespace hdg
{
  // Multifrontal arithmetic of a plan in plain host loops, in the order the
  // device solver runs it, on a random diagonally dominant matrix in the
  // trace pattern; compared with Gaussian elimination of the assembled
  // matrix.  Small meshes only (the dense check is O(ndof^3)).
  bool nd_plan_selftest(int d, const int n[3], int tdf, unsigned seed, const double *given,
                        double out[3], std::string &err)
  {
    Topology T;
    build_topology(d, n, T);
    NDPlan P;
    if (!build_nd_plan(T, tdf, P, err))
      return false;
    const long long N = P.ndof;
    if (N > 3000)
      {
        err = "self-test: mesh too large for the dense check";
        return false;
      }
    unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
    auto rnd = [&]() {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    std::vector<double> vals(P.nnz), A((size_t)N * N, 0.0), rhs(N);
    for (long long f = 0; f < T.nfaces; ++f)
      for (long long bb = T.blkptr[f]; bb < T.blkptr[f + 1]; ++bb)
        for (int r = 0; r < tdf; ++r)
          for (int c = 0; c < tdf; ++c)
            {
              const long long k = T.blkptr[f] * tdf * tdf + (long long)r * T.nbk[f] * tdf +
                                  (bb - T.blkptr[f]) * tdf + c;
              const long long row = f * tdf + r, col = T.blk_cols[bb] * tdf + c;
              double          v   = rnd();
              if (row == col)
                v += 2.0 * tdf * 7;
              if (given)
                v = given[k];
              vals[k]                     = v;
              A[(size_t)row * N + col]    = v;
            }
    for (auto &v : rhs)
      v = rnd();
    // dense reference
    std::vector<double> M = A, xr = rhs;
    for (long long k = 0; k < N; ++k)
      {
        long long p = k;
        for (long long i = k + 1; i < N; ++i)
          if (std::abs(M[i * N + k]) > std::abs(M[p * N + k]))
            p = i;
        if (p != k)
          {
            for (long long j = 0; j < N; ++j)
              std::swap(M[k * N + j], M[p * N + j]);
            std::swap(xr[k], xr[p]);
          }
        for (long long i = k + 1; i < N; ++i)
          {
            const double l = M[i * N + k] / M[k * N + k];
            for (long long j = k; j < N; ++j)
              M[i * N + j] -= l * M[k * N + j];
            xr[i] -= l * xr[k];
          }
      }
    for (long long k = N - 1; k >= 0; --k)
      {
        double v = xr[k];
        for (long long j = k + 1; j < N; ++j)
          v -= M[k * N + j] * xr[j];
        xr[k] = v / M[k * N + k];
      }
    // multifrontal factorisation
    std::vector<double> F(P.Ftot, 0.0), w(P.wtot, 0.0), x(N, 0.0);
    std::vector<int>    ipiv(std::max<long long>(1, P.ptot));
    for (long long k = 0; k < P.nnz; ++k)
      F[P.map[k]] = vals[k];
    for (long long v : P.pad)
      F[v] = 1.0;
    for (const NDGroup &G : P.groups)
      {
        const int S = G.S, I = G.I, n2 = G.n;
        for (int b = 0; b < G.batch; ++b)
          {
            double *Fb = F.data() + G.fbase + (long long)b * n2 * n2;
            int    *pv = ipiv.data() + G.pbase + (long long)b * S;
            auto    at = [&](int i, int j) -> double & { return Fb[i + (long long)j * n2]; };
            for (int k = 0; k < S; ++k)
              {
                int p = k;
                for (int i = k + 1; i < S; ++i)
                  if (std::abs(at(i, k)) > std::abs(at(p, k)))
                    p = i;
                if (at(p, k) == 0.0)
                  {
                    err = "self-test: zero pivot at height " +
                          std::to_string(&G - P.groups.data()) + ", node " + std::to_string(b) +
                          ", column " + std::to_string(k);
                    return false;
                  }
                pv[k] = p + 1;
                if (p != k)
                  for (int j = 0; j < n2; ++j)
                    std::swap(at(k, j), at(p, j));
                for (int i = k + 1; i < S; ++i)
                  {
                    at(i, k) /= at(k, k);
                    for (int j = k + 1; j < n2; ++j)
                      at(i, j) -= at(i, k) * at(k, j);
                  }
              }
            for (int i = S; i < n2; ++i)
              for (int j = 0; j < S; ++j)
                {
                  double v = at(i, j);
                  for (int l = 0; l < j; ++l)
                    v -= at(i, l) * at(l, j);
                  at(i, j) = v / at(j, j);
                }
            for (int i = S; i < n2; ++i)
              for (int j = S; j < n2; ++j)
                for (int l = 0; l < S; ++l)
                  at(i, j) -= at(i, l) * at(l, j);
          }
        for (int want = 0; want < 2; ++want)
          for (int b = 0; b < G.batch; ++b)
            {
              if (G.slot[b] != want || G.pf[b] < 0)
                continue;
              const double *U = F.data() + G.fbase + (long long)b * n2 * n2 + S + (long long)S * n2;
              for (int c = 0; c < I; ++c)
                for (int r = 0; r < I; ++r)
                  {
                    const int pr = G.pos[(size_t)b * I + r], pc = G.pos[(size_t)b * I + c];
                    if (pr >= 0 && pc >= 0)
                      F[G.pf[b] + pr + (long long)pc * G.pn[b]] += U[r + (long long)c * n2];
                  }
            }
      }
    // solve
    for (long long i = 0; i < N; ++i)
      w[P.wmap[i]] = rhs[i];
    for (const NDGroup &G : P.groups)
      {
        const int S = G.S, I = G.I, n2 = G.n;
        for (int b = 0; b < G.batch; ++b)
          {
            const double *Fb = F.data() + G.fbase + (long long)b * n2 * n2;
            double       *wb = w.data() + G.wbase + (long long)b * n2;
            const int    *pv = ipiv.data() + G.pbase + (long long)b * S;
            for (int j = 0; j < S; ++j)
              if (pv[j] - 1 != j)
                std::swap(wb[j], wb[pv[j] - 1]);
            for (int i = 0; i < S; ++i)
              for (int l = 0; l < i; ++l)
                wb[i] -= Fb[i + (long long)l * n2] * wb[l];
            for (int i = S; i < n2; ++i)
              for (int l = 0; l < S; ++l)
                wb[i] -= Fb[i + (long long)l * n2] * wb[l];
          }
        for (int want = 0; want < 2; ++want)
          for (int b = 0; b < G.batch; ++b)
            {
              if (G.slot[b] != want || G.pf[b] < 0)
                continue;
              for (int r = 0; r < I; ++r)
                if (G.pos[(size_t)b * I + r] >= 0)
                  w[G.pw[b] + G.pos[(size_t)b * I + r]] += w[G.wbase + (long long)b * n2 + S + r];
            }
      }
    for (int gi = (int)P.groups.size() - 1; gi >= 0; --gi)
      {
        const NDGroup &G = P.groups[gi];
        const int      S = G.S, I = G.I, n2 = G.n;
        for (int b = 0; b < G.batch; ++b)
          {
            const double *Fb = F.data() + G.fbase + (long long)b * n2 * n2;
            double       *wb = w.data() + G.wbase + (long long)b * n2;
            for (int r = 0; r < I; ++r)
              {
                const long long dof = G.idof[(size_t)b * I + r];
                wb[S + r]           = dof >= 0 ? x[dof] : 0.0;
              }
            for (int i = 0; i < S; ++i)
              for (int r = 0; r < I; ++r)
                wb[i] -= Fb[i + (long long)(S + r) * n2] * wb[S + r];
            for (int i = S - 1; i >= 0; --i)
              {
                for (int l = i + 1; l < S; ++l)
                  wb[i] -= Fb[i + (long long)l * n2] * wb[l];
                wb[i] /= Fb[i + (long long)i * n2];
              }
            for (int j = 0; j < S; ++j)
              {
                const long long dof = G.edof[(size_t)b * S + j];
                if (dof >= 0)
                  x[dof] = wb[j];
              }
          }
      }
    double num = 0, den = 0;
    for (long long i = 0; i < N; ++i)
      {
        num += (x[i] - xr[i]) * (x[i] - xr[i]);
        den += xr[i] * xr[i];
      }
    out[0] = std::sqrt(num / den);
    out[1] = (double)P.nnodes;
    out[2] = (double)P.groups.size();
    return true;
  }
} // namespace hdg
