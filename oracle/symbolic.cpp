// Oracle (test infrastructure only): host symbolic phase restated from the
// reference.  Deliberately simple loops; the product library has its own,
// faster implementations that must agree with these bit for bit.
//
//   orc_node_graph      bsr_utils.h:430-476 get_row_col_ptrs (node adjacency
//                       incl. self, sorted unique), here via per-node element
//                       lists instead of the O(nnodes*nelems) scan
//   orc_rcm             reordering_utils.h:189-291 TacsComputeRCMLevSetOrder /
//                       TacsComputeRCMOrder (rcm_vars[old] = new)
//   orc_permute_pattern gmres.cu:242-252 BSRMatApplyPerm: B(perm[i],perm[j])
//                       nonzero iff A(i,j) nonzero, sorted columns
//   orc_iluk            bsr_utils.h:678-807 computeILUk (level-of-fill sum rule)
//   orc_elem_ind_map    bsr_utils.h:809-870 get_elem_ind_map
//   orc_transpose       bsr_utils.h:157-211 get_transpose_bsr_data
//   orc_bandwidth       bsr_utils.h:655-670 getBandWidth
//   orc_full_fill       symbolic LU fill (ILU(k) with k = infinity)
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <vector>

#include "oracle.h"

namespace {
int* copy_out(const std::vector<int>& v) {
  int* p = (int*)std::malloc(sizeof(int) * (v.size() ? v.size() : 1));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(int) * v.size());
  return p;
}
}  // namespace

extern "C" void orc_free(void* p) { std::free(p); }

extern "C" int orc_node_graph(int nelems, int nnodes, int npe, const int* conn,
                              int** rowp, int** cols) {
  std::vector<std::vector<int>> elems_of(nnodes);
  for (int e = 0; e < nelems; e++)
    for (int a = 0; a < npe; a++) elems_of[conn[npe * e + a]].push_back(e);
  std::vector<int> rp(nnodes + 1, 0), cl;
  for (int i = 0; i < nnodes; i++) {
    std::vector<int> tmp;
    for (int e : elems_of[i])
      for (int a = 0; a < npe; a++) tmp.push_back(conn[npe * e + a]);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    cl.insert(cl.end(), tmp.begin(), tmp.end());
    rp[i + 1] = (int)cl.size();
  }
  *rowp = copy_out(rp);
  *cols = copy_out(cl);
  return rp[nnodes];
}

extern "C" int orc_rcm(int n, const int* rowp, const int* cols, int root,
                       int* rcm_vars) {
  // breadth-first level sets from root, neighbours in stored (ascending)
  // order; a new component starts at the lowest unnumbered node; the final
  // numbering is reversed.
  std::vector<int> levset(n);
  for (int k = 0; k < n; k++) rcm_vars[k] = -1;
  int start = 0, end = 0, var_num = 0;
  while (var_num < n) {
    if (end - start == 0) {
      if (rcm_vars[root] >= 0) {
        int i = 0;
        for (; i < n; i++)
          if (rcm_vars[i] < 0) {
            root = i;
            break;
          }
        if (i >= n) return var_num;
      }
      levset[end] = root;
      rcm_vars[root] = var_num++;
      end++;
    }
    while (start < end) {
      int next = end;
      for (int cur = start; cur < end; cur++) {
        int node = levset[cur];
        for (int j = rowp[node]; j < rowp[node + 1]; j++) {
          int nb = cols[j];
          if (rcm_vars[nb] < 0) {
            rcm_vars[nb] = var_num++;
            levset[next++] = nb;
          }
        }
      }
      start = end;
      end = next;
    }
  }
  for (int i = 0; i < var_num; i++) rcm_vars[levset[i]] = var_num - 1 - i;
  return var_num;
}

extern "C" int orc_permute_pattern(int n, const int* rowp, const int* cols,
                                   const int* perm, int** prowp, int** pcols) {
  std::vector<int> iperm(n);
  for (int i = 0; i < n; i++) iperm[perm[i]] = i;
  std::vector<int> rp(n + 1, 0), cl;
  for (int r = 0; r < n; r++) {
    int old = iperm[r];
    std::vector<int> tmp;
    for (int j = rowp[old]; j < rowp[old + 1]; j++) tmp.push_back(perm[cols[j]]);
    std::sort(tmp.begin(), tmp.end());
    cl.insert(cl.end(), tmp.begin(), tmp.end());
    rp[r + 1] = (int)cl.size();
  }
  *prowp = copy_out(rp);
  *pcols = copy_out(cl);
  return rp[n];
}

extern "C" int orc_iluk(int n, const int* rowp, const int* cols, int lev_fill,
                        int** frowp, int** fcols, int** flevs) {
  // Row-by-row symbolic ILU(k): a candidate (i, c) reached through pivot k
  // gets level lev(i,k) + lev(k,c) + 1; existing entries keep the minimum;
  // new entries are admitted when the level is <= lev_fill; a missing
  // diagonal is inserted.  Pivots are visited in ascending column order,
  // including fill created while the row is being processed.
  std::vector<int> rp(n + 1, 0), cl, lv, diag(n);
  std::vector<int> rcols, rlevs;
  for (int i = 0; i < n; i++) {
    rcols.assign(cols + rowp[i], cols + rowp[i + 1]);
    rlevs.assign(rcols.size(), 0);
    if (!std::binary_search(rcols.begin(), rcols.end(), i)) {
      auto it = std::lower_bound(rcols.begin(), rcols.end(), i);
      size_t pos = it - rcols.begin();
      rcols.insert(rcols.begin() + pos, i);
      rlevs.insert(rlevs.begin() + pos, 0);
    }
    size_t j = 0;
    for (; rcols[j] < i; j++) {
      int k = rcols[j];
      int clev = rlevs[j];
      size_t p = j + 1;
      for (int kk = diag[k] + 1; kk < rp[k + 1]; kk++) {
        int c = cl[kk];
        int nl = clev + lv[kk] + 1;
        while (p < rcols.size() && rcols[p] < c) p++;
        if (p < rcols.size() && rcols[p] == c) {
          if (rlevs[p] > nl) rlevs[p] = nl;
        } else if (nl <= lev_fill) {
          rcols.insert(rcols.begin() + p, c);
          rlevs.insert(rlevs.begin() + p, nl);
        }
      }
    }
    diag[i] = (int)cl.size() + (int)j;
    cl.insert(cl.end(), rcols.begin(), rcols.end());
    lv.insert(lv.end(), rlevs.begin(), rlevs.end());
    rp[i + 1] = (int)cl.size();
  }
  *frowp = copy_out(rp);
  *fcols = copy_out(cl);
  if (flevs) *flevs = copy_out(lv);
  return rp[n];
}

extern "C" int orc_full_fill(int n, const int* rowp, const int* cols,
                             int** frowp, int** fcols) {
  return orc_iluk(n, rowp, cols, 1 << 29, frowp, fcols, nullptr);
}

extern "C" void orc_elem_ind_map(int nelems, int npe, const int* conn,
                                 const int* rowp, const int* cols,
                                 const int* perm, int* map) {
  int npe2 = npe * npe;
  for (int e = 0; e < nelems; e++) {
    const int* lc = &conn[npe * e];
    for (int a = 0; a < npe; a++) {
      int row = perm[lc[a]];
      for (int b = 0; b < npe; b++) {
        int col = perm[lc[b]];
        map[npe2 * e + npe * a + b] = 0;  // reference leaves 0 if absent
        for (int i = rowp[row]; i < rowp[row + 1]; i++)
          if (cols[i] == col) {
            map[npe2 * e + npe * a + b] = i;
            break;
          }
      }
    }
  }
}

extern "C" int orc_transpose(int n, const int* rowp, const int* cols,
                             int** trowp, int** tcols, int** tmap) {
  int nnz = rowp[n];
  std::vector<int> cnt(n + 1, 0);
  for (int k = 0; k < nnz; k++) cnt[cols[k] + 1]++;
  for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
  std::vector<int> tr(cnt), tc(nnz), tm(nnz), fill(n, 0);
  for (int r = 0; r < n; r++)
    for (int k = rowp[r]; k < rowp[r + 1]; k++) {
      int c = cols[k];
      int t = tr[c] + fill[c]++;
      tc[t] = r;
      tm[t] = k;
    }
  *trowp = copy_out(tr);
  *tcols = copy_out(tc);
  *tmap = copy_out(tm);
  return nnz;
}

extern "C" int orc_bandwidth(int n, const int* rowp, const int* cols) {
  int bw = 0;
  for (int i = 0; i < n; i++)
    for (int jp = rowp[i]; jp < rowp[i + 1]; jp++)
      bw = std::max(bw, std::abs(i - cols[jp]));
  return bw;
}

// Exact minimum-degree ordering on the elimination graph (tiny inputs only):
// the checker for the product's approximate-minimum-degree ordering uses
// invariants (permutation, fill no worse than a bound), so this is a plain
// quadratic restatement of the greedy rule with lowest-index tie break.
extern "C" int orc_min_degree(int n, const int* rowp, const int* cols,
                              int* perm) {
  std::vector<std::vector<char>> adj(n, std::vector<char>(n, 0));
  for (int i = 0; i < n; i++)
    for (int k = rowp[i]; k < rowp[i + 1]; k++)
      if (cols[k] != i) adj[i][cols[k]] = adj[cols[k]][i] = 1;
  std::vector<char> done(n, 0);
  for (int step = 0; step < n; step++) {
    int best = -1, bdeg = 1 << 30;
    for (int i = 0; i < n; i++) {
      if (done[i]) continue;
      int d = 0;
      for (int j = 0; j < n; j++) d += (!done[j] && adj[i][j]);
      if (d < bdeg) {
        bdeg = d;
        best = i;
      }
    }
    perm[best] = step;
    done[best] = 1;
    for (int a = 0; a < n; a++)
      if (!done[a] && adj[best][a])
        for (int b = 0; b < n; b++)
          if (b != a && !done[b] && adj[best][b]) adj[a][b] = adj[b][a] = 1;
  }
  return n;
}

// Minimum-degree checks on the explicit elimination graph (dense bitsets;
// small graphs only).  The reference takes its AMD ordering from the absent
// sparse-utils library, so an ordering is checked against what minimum
// degree means (Rose's elimination-graph model; Tinney-Walker scheme 2)
// rather than against a reference permutation.
namespace {
struct ElimGraph {
  int n, words;
  std::vector<uint64_t> adj, alive;
  ElimGraph(int n_, const int* rowp, const int* cols)
      : n(n_), words((n_ + 63) / 64), adj((size_t)n_ * ((n_ + 63) / 64), 0),
        alive((n_ + 63) / 64, 0) {
    for (int i = 0; i < n; i++) {
      alive[i / 64] |= 1ull << (i % 64);
      for (int k = rowp[i]; k < rowp[i + 1]; k++) {
        const int j = cols[k];
        if (j == i) continue;
        row(i)[j / 64] |= 1ull << (j % 64);
        row(j)[i / 64] |= 1ull << (i % 64);
      }
    }
  }
  uint64_t* row(int i) { return &adj[(size_t)i * words]; }
  // number of uneliminated neighbours
  int degree(int i) {
    int d = 0;
    for (int w = 0; w < words; w++) d += __builtin_popcountll(row(i)[w] & alive[w]);
    return d;
  }
  // eliminates p: its uneliminated neighbours become a clique; returns them
  std::vector<int> eliminate(int p) {
    alive[p / 64] &= ~(1ull << (p % 64));
    std::vector<uint64_t> nb(words);
    std::vector<int> out;
    for (int w = 0; w < words; w++) nb[w] = row(p)[w] & alive[w];
    for (int w = 0; w < words; w++)
      for (uint64_t b = nb[w]; b; b &= b - 1) out.push_back(64 * w + __builtin_ctzll(b));
    for (int a : out) {
      for (int w = 0; w < words; w++) row(a)[w] |= nb[w];
      row(a)[a / 64] &= ~(1ull << (a % 64));
    }
    return out;
  }
};
}  // namespace

// Replays perm (perm[old] = new): dpiv[k] = exact degree of the k-th pivot
// when it is eliminated, dmin[k] = the least exact degree of any node left
// at that step.  Returns the blocks of the complete factor's lower triangle
// including the diagonal, sum_k (dpiv[k] + 1).
extern "C" long long orc_md_replay(int n, const int* rowp, const int* cols,
                                   const int* perm, int* dpiv, int* dmin) {
  ElimGraph g(n, rowp, cols);
  std::vector<int> iperm(n);
  for (int i = 0; i < n; i++) iperm[perm[i]] = i;
  long long nnz = 0;
  for (int k = 0; k < n; k++) {
    int best = n;
    for (int i = 0; i < n; i++)
      if (g.alive[i / 64] >> (i % 64) & 1) best = std::min(best, g.degree(i));
    const int p = iperm[k];
    const int d = g.degree(p);
    if (dpiv) dpiv[k] = d;
    if (dmin) dmin[k] = best;
    nnz += d + 1;
    g.eliminate(p);
  }
  return nnz;
}

// Exact minimum degree with the tie-break of a bucket-stack implementation:
// among the nodes of least exact degree, the one whose degree was updated
// most recently; nodes updated by the same elimination in descending index
// order, nodes never updated in ascending index order (the buckets are
// filled from the highest index down, and an update pushes a node on top of
// its bucket).  Every neighbour of an eliminated node counts as updated.
extern "C" int orc_min_degree_stack(int n, const int* rowp, const int* cols,
                                    int* perm) {
  ElimGraph g(n, rowp, cols);
  std::vector<long long> stamp(n);
  for (int i = 0; i < n; i++) stamp[i] = n - 1 - i;
  long long clock = n;
  for (int k = 0; k < n; k++) {
    int best = -1, bd = 0;
    for (int i = 0; i < n; i++) {
      if (!(g.alive[i / 64] >> (i % 64) & 1)) continue;
      const int d = g.degree(i);
      if (best < 0 || d < bd || (d == bd && stamp[i] > stamp[best])) {
        best = i;
        bd = d;
      }
    }
    perm[best] = k;
    for (int a : g.eliminate(best)) stamp[a] = clock++;
  }
  return n;
}
