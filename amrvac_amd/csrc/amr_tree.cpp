// amr_tree.cpp -- native host routines of the block-structured AMR tree:
// the neighbour tables of the leaves (build_connectivity) and the 2:1
// balance of the regrid flags (proper_nesting).  Plain C++ (no device code).
#include "amrvac_amr.h"

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

constexpr int NB_BOUNDARY = 1, NB_COARSE = 2, NB_SIBLING = 3, NB_FINE = 4;
constexpr int kInternal = -1;  // map value of a node that is not a leaf

// floor division by 2 (indices can be 0 or negative next to the domain)
inline int fdiv2(int a) { return a >= 0 ? a / 2 : -((-a + 1) / 2); }

struct Tree {
  int nd;
  int ng1[3];
  int periodic[3];
  std::unordered_map<uint64_t, int> node;  // (level, ig) -> leaf block, or kInternal

  static uint64_t key(int level, const int* ig) {
    return (uint64_t)level << 60 | (uint64_t)ig[0] << 40 | (uint64_t)ig[1] << 20 | (uint64_t)ig[2];
  }

  int build(int nleaves, const int* level, const int* ig) {
    node.clear();
    node.reserve((size_t)nleaves * 2);
    for (int b = 0; b < nleaves; ++b) {
      int g[3] = {1, 1, 1};
      for (int d = 0; d < nd; ++d) g[d] = ig[b * nd + d];
      const int lev = level[b];
      if (lev < 1 || lev > 15) return fail(1, "level out of range (1..15)");
      for (int d = 0; d < nd; ++d)
        if (g[d] < 1 || g[d] > (ng1[d] << (lev - 1)) || g[d] >= (1 << 20))
          return fail(1, "block index out of range");
      auto ins = node.emplace(key(lev, g), b);
      if (!ins.second) {
        if (ins.first->second != kInternal) return fail(1, "two leaves at one position");
        return fail(1, "a leaf lies below another leaf");
      }
      // ancestors, up to the first one already present
      for (int l = lev - 1; l >= 1; --l) {
        for (int d = 0; d < nd; ++d) g[d] = (g[d] - 1) / 2 + 1;
        auto a = node.emplace(key(l, g), kInternal);
        if (!a.second) {
          if (a.first->second != kInternal) return fail(1, "a leaf lies below another leaf");
          break;
        }
      }
    }
    return 0;
  }

  enum Status { OUT, NONE, FOUND };

  // the node at (level, ig) with periodic wrap; g receives the wrapped index
  Status at(int level, const int* ig, int* g, int& value) const {
    g[0] = g[1] = g[2] = 1;
    for (int d = 0; d < nd; ++d) {
      const int n = ng1[d] << (level - 1);
      int v = ig[d];
      if (v < 1 || v > n) {
        if (!periodic[d]) return OUT;
        v = 1 + (((v - 1) % n) + n) % n;
      }
      g[d] = v;
    }
    auto it = node.find(key(level, g));
    if (it == node.end()) return NONE;
    value = it->second;
    return FOUND;
  }

  // find_neighbor: type, the neighbour's map value (leaf block or
  // kInternal) and its wrapped index
  int neighbor(int level, const int* ig, const int* i, int& value, int* g) const {
    int tgt[3] = {1, 1, 1};
    for (int d = 0; d < nd; ++d) tgt[d] = ig[d] + i[d];
    const Status s = at(level, tgt, g, value);
    if (s == OUT) return NB_BOUNDARY;
    if (s == NONE) {
      // no node at this level: the parent's neighbour is a leaf
      int pt[3] = {1, 1, 1};
      for (int d = 0; d < nd; ++d) pt[d] = fdiv2(tgt[d] - 1) + 1;
      if (level < 2 || at(level - 1, pt, g, value) != FOUND) return NB_BOUNDARY;
      return NB_COARSE;
    }
    return value == kInternal ? NB_FINE : NB_SIBLING;
  }

  int child_of(int level, const int* g, int j, int* cg) const {
    int v = kInternal;
    int c[3] = {1, 1, 1};
    for (int d = 0; d < nd; ++d) c[d] = 2 * (g[d] - 1) + 1 + ((j >> d) & 1);
    for (int d = 0; d < 3; ++d) cg[d] = c[d];
    auto it = node.find(key(level + 1, c));
    if (it == node.end()) return -2;  // no such node
    v = it->second;
    return v;
  }
};

int pow3(int nd) { return nd == 1 ? 3 : nd == 2 ? 9 : 27; }

bool init_tree(Tree& t, int ndim, const int* ng1, const int* periodic) {
  if (ndim < 1 || ndim > 3) return false;
  t.nd = ndim;
  for (int d = 0; d < 3; ++d) {
    t.ng1[d] = d < ndim ? ng1[d] : 1;
    t.periodic[d] = d < ndim ? periodic[d] : 0;
  }
  return true;
}

}  // namespace

extern "C" {

const char* amrvac_amr_last_error(void) { return g_err.c_str(); }

int amrvac_amr_connectivity(int ndim, const int* ng1, const int* periodic, int nleaves, const int* level,
                            const int* ig, int* ntype, int* nbr, int* nchild) {
  Tree t;
  if (!init_tree(t, ndim, ng1, periodic)) return fail(1, "ndim must be 1, 2 or 3");
  if (int rc = t.build(nleaves, level, ig)) return rc;
  const int nn = pow3(ndim), nc = 1 << ndim;
  for (int b = 0; b < nleaves; ++b) {
    const int* gb = ig + (size_t)b * ndim;
    for (int k = 0; k < nn; ++k) {
      int i[3] = {0, 0, 0};
      for (int d = 0, r = k; d < ndim; ++d, r /= 3) i[d] = r % 3 - 1;
      const size_t o = (size_t)b * nn + k;
      for (int j = 0; j < nc; ++j) nchild[o * nc + j] = -1;
      nbr[o] = -1;
      if (i[0] == 0 && i[1] == 0 && i[2] == 0) {
        ntype[o] = 0;
        nbr[o] = b;
        continue;
      }
      int value = kInternal, g[3];
      const int ty = t.neighbor(level[b], gb, i, value, g);
      ntype[o] = ty;
      if (ty == NB_SIBLING || ty == NB_COARSE) {
        nbr[o] = value;
      } else if (ty == NB_FINE) {
        // the children of the neighbour node that touch this block
        for (int j = 0; j < nc; ++j) {
          bool touch = true;
          for (int d = 0; d < ndim; ++d) {
            const int icj = 1 + ((j >> d) & 1);
            if ((i[d] == -1 && icj != 2) || (i[d] == 1 && icj != 1)) touch = false;
          }
          if (!touch) continue;
          int cg[3];
          const int cv = t.child_of(level[b], g, j, cg);
          if (cv < 0) return fail(2, "level jump larger than one between neighbours");
          nchild[o * nc + j] = cv;
        }
      }
    }
  }
  return 0;
}

int amrvac_amr_proper_nesting(int ndim, const int* ng1, const int* periodic, int nleaves, const int* level,
                              const int* ig, const int* order, int refine_max_level, unsigned char* refine,
                              unsigned char* coarsen) {
  Tree t;
  if (!init_tree(t, ndim, ng1, periodic)) return fail(1, "ndim must be 1, 2 or 3");
  if (int rc = t.build(nleaves, level, ig)) return rc;
  if (nleaves == 0) return 0;
  int levmin = level[0], levmax = level[0];
  for (int b = 1; b < nleaves; ++b) {
    levmin = level[b] < levmin ? level[b] : levmin;
    levmax = level[b] > levmax ? level[b] : levmax;
  }
  const int nc = 1 << ndim;
  // outward directions of a child at position ic: i_d in {ic_d - 2, ic_d - 1},
  // not all zero
  auto outward = [&](const int* g, auto&& fn) {
    int lo[3] = {0, 0, 0}, n[3] = {1, 1, 1};
    for (int d = 0; d < ndim; ++d) {
      lo[d] = (1 + (g[d] - 1) % 2) - 2;
      n[d] = 2;
    }
    for (int a2 = 0; a2 < n[2]; ++a2)
      for (int a1 = 0; a1 < n[1]; ++a1)
        for (int a0 = 0; a0 < n[0]; ++a0) {
          const int i[3] = {lo[0] + a0, lo[1] + a1, lo[2] + a2};
          if (i[0] == 0 && i[1] == 0 && i[2] == 0) continue;
          if (fn(i)) return true;
        }
    return false;
  };
  // refinement spreads to the leaves next to the parent, coarser levels last
  const int top = levmax < refine_max_level - 1 ? levmax : refine_max_level - 1;
  for (int lev = top; lev > levmin; --lev)
    for (int b = 0; b < nleaves; ++b) {
      if (level[b] != lev || !refine[b]) continue;
      const int* gb = ig + (size_t)b * ndim;
      int p[3] = {1, 1, 1};
      for (int d = 0; d < ndim; ++d) p[d] = (gb[d] - 1) / 2 + 1;
      outward(gb, [&](const int* i) {
        int tg[3] = {1, 1, 1}, g[3], v = kInternal;
        for (int d = 0; d < ndim; ++d) tg[d] = p[d] + i[d];
        if (t.at(lev - 1, tg, g, v) == Tree::FOUND && v != kInternal) refine[v] = 1;
        return false;
      });
    }
  for (int b = 0; b < nleaves; ++b)
    if (refine[b]) coarsen[b] = 0;
  // coarsening: whole families, checked in the level lists' order
  const int bottom = levmin > 2 ? levmin : 2;
  for (int lev = levmax; lev >= bottom; --lev)
    for (int q = 0; q < nleaves; ++q) {
      const int b = order[q];
      if (b < 0 || b >= nleaves) return fail(1, "order entry out of range");
      if (level[b] != lev || !coarsen[b]) continue;
      const int* gb = ig + (size_t)b * ndim;
      int pg[3] = {1, 1, 1};
      for (int d = 0; d < ndim; ++d) pg[d] = (gb[d] - 1) / 2 + 1;
      int sib[8];
      int sg[8][3];
      bool all = true;
      for (int j = 0; j < nc; ++j) {
        sib[j] = t.child_of(lev - 1, pg, j, sg[j]);
        if (sib[j] < 0 || !coarsen[sib[j]]) all = false;
      }
      auto unflag = [&]() {
        for (int j = 0; j < nc; ++j)
          if (sib[j] >= 0) coarsen[sib[j]] = 0;
      };
      if (!all) {
        unflag();
        continue;
      }
      bool ok = true;
      for (int j = 0; j < nc && ok; ++j) {
        const bool bad = outward(sg[j], [&](const int* i) {
          int v = kInternal, g[3];
          const int ty = t.neighbor(lev, sg[j], i, v, g);
          if (ty == NB_SIBLING) return refine[v] != 0;
          if (ty == NB_FINE) {
            int cg[3];
            const int c1 = t.child_of(lev, g, 0, cg);
            return !(c1 >= 0 && coarsen[c1]);
          }
          return false;
        });
        if (bad) ok = false;
      }
      if (!ok) unflag();
    }
  return 0;
}

}  // extern "C"
