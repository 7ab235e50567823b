// Host-side setup of the solve path (native C++, runs once per mesh):
//   * reverse Cuthill-McKee ordering of a node graph,
//   * the node graph / 3x3 block pattern of a C3D10 mesh with the Dirichlet
//     pattern rule,
//   * the per-block contribution lists that make GPU assembly a deterministic
//     gather, and the node -> (element, local node) incidence lists.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace wn {

// Reverse Cuthill-McKee with the semantics of scipy.sparse.csgraph's
// reverse_cuthill_mckee (symmetric_mode=True), the routine the reference calls
// at wnfea/mesh/rcm.py:85:
//   * degree(i) = number of stored entries of row i, plus one more when the
//     diagonal entry is stored;
//   * components are seeded in the order `seed_order` (numpy's argsort of the
//     degrees in the reference), skipping visited nodes;
//   * breadth-first levels: the unvisited neighbours of each node, in column
//     order, are appended and that node's new batch is insertion-sorted by
//     degree (stable for equal degrees);
//   * the final order is reversed.
// Writes perm[new] = old.
void rcm_order(int n, const int* ptr, const int* ind, const int* seed_order, int* perm) {
  std::vector<int> deg(n), order(n);
  std::vector<char> seen(n, 0);
  for (int i = 0; i < n; ++i) {
    int d = ptr[i + 1] - ptr[i];
    for (int k = ptr[i]; k < ptr[i + 1]; ++k)
      if (ind[k] == i) {
        ++d;
        break;
      }
    deg[i] = d;
  }
  int filled = 0;
  for (int zz = 0; zz < n && filled < n; ++zz) {
    const int seed = seed_order[zz];
    if (seen[seed]) continue;
    seen[seed] = 1;
    order[filled++] = seed;
    int head = filled - 1, level_end = filled;
    while (head < level_end) {
      for (int h = head; h < level_end; ++h) {
        const int v = order[h];
        const int batch = filled;
        for (int k = ptr[v]; k < ptr[v + 1]; ++k) {
          const int w = ind[k];
          if (!seen[w]) {
            seen[w] = 1;
            order[filled++] = w;
          }
        }
        // stable insertion sort of the batch by degree
        for (int a = batch + 1; a < filled; ++a) {
          const int node = order[a], d = deg[node];
          int b = a;
          while (b > batch && d < deg[order[b - 1]]) {
            order[b] = order[b - 1];
            --b;
          }
          order[b] = node;
        }
      }
      head = level_end;
      level_end = filled;
    }
  }
  for (int i = 0; i < n; ++i) perm[i] = order[n - 1 - i];
}

// Node graph of a C3D10 mesh: row i lists (sorted, unique) every node sharing
// an element with i, including i.  With `node_free` given, the off-diagonal
// pair (i, j) is kept only when both nodes have at least one free DOF: the
// Dirichlet rule of the reference zeroes every entry of a block whose row node
// or column node is fully fixed, and eliminate_zeros then drops it
// (solid_solver.py:95-127, bsr_matrix.py:84-86).  Returns the number of
// entries, or -1 when `cap` is too small.
long long node_graph(int n_nodes, int M, int npe, const int* conn, const unsigned char* node_free, int* rowptr,
                     int* colidx, long long cap) {
  // node -> element incidence
  std::vector<int> cnt(n_nodes + 1, 0);
  for (long long q = 0; q < (long long)M * npe; ++q) cnt[conn[q] + 1]++;
  for (int i = 0; i < n_nodes; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> inc(cnt[n_nodes]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int e = 0; e < M; ++e)
    for (int a = 0; a < npe; ++a) inc[fill[conn[(long long)e * npe + a]]++] = e;
  std::vector<int> mark(n_nodes, -1), buf;
  long long nnz = 0;
  rowptr[0] = 0;
  for (int i = 0; i < n_nodes; ++i) {
    buf.clear();
    const bool fi = !node_free || node_free[i];
    for (int k = cnt[i]; k < cnt[i + 1]; ++k) {
      const int e = inc[k];
      for (int a = 0; a < npe; ++a) {
        const int j = conn[(long long)e * npe + a];
        if (mark[j] == i) continue;
        mark[j] = i;
        if (j != i && !(fi && (!node_free || node_free[j]))) continue;
        buf.push_back(j);
      }
    }
    if (mark[i] != i) buf.push_back(i);  // isolated node: keep its diagonal block
    std::sort(buf.begin(), buf.end());
    if (nnz + (long long)buf.size() > cap) return -1;
    std::memcpy(colidx + nnz, buf.data(), sizeof(int) * buf.size());
    nnz += buf.size();
    rowptr[i + 1] = static_cast<int>(nnz);
  }
  return nnz;
}

// Contribution lists of a block pattern: for block k = (i, j), the codes
// e*npe*npe + a*npe + b of every element e with conn[e][a] = i and
// conn[e][b] = j, in element order then (a, b) order -- the order in which the
// reference's COO->CSR conversion meets duplicates (solid_solver.py:62-84).
// Pairs whose block is absent from the pattern (Dirichlet-dropped) are skipped.
// cptr has nnzb+1 entries; returns the number of codes or -1 if cap too small.
long long contributions(int n_nodes, int M, int npe, const int* conn, const int* rowptr, const int* colidx,
                        long long* cptr, int* codes, long long cap) {
  const long long nnzb = rowptr[n_nodes];
  std::vector<long long> slot((size_t)M * npe * npe, -1);
  std::vector<long long> count(nnzb + 1, 0);
  for (int e = 0; e < M; ++e) {
    const int* en = conn + (long long)e * npe;
    for (int a = 0; a < npe; ++a) {
      const int i = en[a];
      const int* b0 = colidx + rowptr[i];
      const int* b1 = colidx + rowptr[i + 1];
      for (int b = 0; b < npe; ++b) {
        const int* p = std::lower_bound(b0, b1, en[b]);
        if (p != b1 && *p == en[b]) {
          const long long k = rowptr[i] + (p - b0);
          slot[((size_t)e * npe + a) * npe + b] = k;
          count[k + 1]++;
        }
      }
    }
  }
  for (long long k = 0; k < nnzb; ++k) count[k + 1] += count[k];
  if (count[nnzb] > cap) return -1;
  std::memcpy(cptr, count.data(), sizeof(long long) * (nnzb + 1));
  std::vector<long long> pos(count.begin(), count.end() - 1);
  for (long long q = 0; q < (long long)M * npe * npe; ++q) {
    const long long k = slot[q];
    if (k >= 0) codes[pos[k]++] = static_cast<int>(q);
  }
  return count[nnzb];
}

// node -> (element*npe + local) incidence in element order.
void node_incidence(int n_nodes, int M, int npe, const int* conn, int* ptr, int* codes) {
  std::memset(ptr, 0, sizeof(int) * (n_nodes + 1));
  for (long long q = 0; q < (long long)M * npe; ++q) ptr[conn[q] + 1]++;
  for (int i = 0; i < n_nodes; ++i) ptr[i + 1] += ptr[i];
  std::vector<int> pos(ptr, ptr + n_nodes);
  for (long long q = 0; q < (long long)M * npe; ++q) codes[pos[conn[q]]++] = static_cast<int>(q);
}

}  // namespace wn

namespace wn {

// Smoothed-aggregation coarsening of a node strength graph (CSR without the
// diagonal).  Standard three phases:
//   1. a node whose strong neighbours are all unaggregated roots a new
//      aggregate made of itself and those neighbours;
//   2. a leftover node joins the phase-1 aggregate most represented among its
//      neighbours (lowest aggregate index on ties);
//   3. remaining nodes root new aggregates with their unaggregated neighbours.
// Nodes without strong neighbours (e.g. fully clamped nodes, whose rows are
// identity) stay unaggregated: agg = -1.  Returns the number of aggregates.
long long sa_aggregate(int n, const int* ptr, const int* ind, int* agg) {
  for (int i = 0; i < n; ++i) agg[i] = -1;
  int na = 0;
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0 || ptr[i + 1] == ptr[i]) continue;
    bool free_nb = true;
    for (int k = ptr[i]; k < ptr[i + 1] && free_nb; ++k) free_nb = agg[ind[k]] < 0;
    if (!free_nb) continue;
    agg[i] = na;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) agg[ind[k]] = na;
    ++na;
  }
  std::vector<int> phase1(agg, agg + n);
  std::vector<int> votes;
  std::vector<int> touched;
  votes.assign(na + 1, 0);
  for (int i = 0; i < n; ++i) {
    if (phase1[i] >= 0 || ptr[i + 1] == ptr[i]) continue;
    touched.clear();
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int a = phase1[ind[k]];
      if (a >= 0) {
        if (votes[a] == 0) touched.push_back(a);
        votes[a]++;
      }
    }
    int best = -1, bv = 0;
    for (int a : touched) {
      if (votes[a] > bv || (votes[a] == bv && a < best)) {
        bv = votes[a];
        best = a;
      }
    }
    for (int a : touched) votes[a] = 0;
    if (best >= 0) agg[i] = best;
  }
  for (int i = 0; i < n; ++i) {
    if (agg[i] >= 0 || ptr[i + 1] == ptr[i]) continue;
    agg[i] = na;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k)
      if (agg[ind[k]] < 0) agg[ind[k]] = na;
    ++na;
  }
  return na;
}

// Tentative prolongator of smoothed aggregation.  B holds the near-null-space
// (n_nodes*b rows x k columns, row-major).  For each aggregate, the rows of its
// nodes (ascending node order) are orthonormalised column by column (modified
// Gram-Schmidt, applied twice); Q becomes that aggregate's block column of P
// and R (k x k, upper triangular) its rows of the coarse near-null-space Bc.
// Columns that vanish (rank-deficient aggregates) are dropped as zero columns.
// P is written as block CSR with b x k blocks: one block per aggregated node,
// none for an unaggregated node.  Returns 0.
int sa_tentative(int n_nodes, int b, int k, const int* agg, int n_agg, const double* B, int* p_rowptr, int* p_colidx,
                 double* p_vals, double* Bc) {
  std::vector<int> cnt(n_agg + 1, 0);
  for (int i = 0; i < n_nodes; ++i)
    if (agg[i] >= 0) cnt[agg[i] + 1]++;
  for (int a = 0; a < n_agg; ++a) cnt[a + 1] += cnt[a];
  std::vector<int> members(cnt[n_agg]);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < n_nodes; ++i)
    if (agg[i] >= 0) members[fill[agg[i]]++] = i;
  // block-row structure of P
  p_rowptr[0] = 0;
  std::vector<long long> blk_of_node(n_nodes, -1);
  for (int i = 0; i < n_nodes; ++i) {
    p_rowptr[i + 1] = p_rowptr[i] + (agg[i] >= 0 ? 1 : 0);
    if (agg[i] >= 0) {
      blk_of_node[i] = p_rowptr[i];
      p_colidx[p_rowptr[i]] = agg[i];
    }
  }
  std::memset(Bc, 0, sizeof(double) * (size_t)n_agg * k * k);
  std::vector<double> Q, R;
  for (int a = 0; a < n_agg; ++a) {
    const int m_nodes = cnt[a + 1] - cnt[a];
    const int m = m_nodes * b;
    Q.assign((size_t)m * k, 0.0);
    R.assign((size_t)k * k, 0.0);
    for (int t = 0; t < m_nodes; ++t) {
      const int node = members[cnt[a] + t];
      for (int r = 0; r < b; ++r)
        for (int c = 0; c < k; ++c) Q[(size_t)(t * b + r) * k + c] = B[((size_t)node * b + r) * k + c];
    }
    for (int j = 0; j < k; ++j) {
      double n0 = 0.0;
      for (int i = 0; i < m; ++i) n0 += Q[(size_t)i * k + j] * Q[(size_t)i * k + j];
      n0 = std::sqrt(n0);
      for (int pass = 0; pass < 2; ++pass)
        for (int p = 0; p < j; ++p) {
          double d = 0.0;
          for (int i = 0; i < m; ++i) d += Q[(size_t)i * k + p] * Q[(size_t)i * k + j];
          R[p * k + j] += d;
          for (int i = 0; i < m; ++i) Q[(size_t)i * k + j] -= d * Q[(size_t)i * k + p];
        }
      double nj = 0.0;
      for (int i = 0; i < m; ++i) nj += Q[(size_t)i * k + j] * Q[(size_t)i * k + j];
      nj = std::sqrt(nj);
      if (nj > 1e-10 * n0 && nj > 0.0) {
        R[j * k + j] = nj;
        for (int i = 0; i < m; ++i) Q[(size_t)i * k + j] /= nj;
      } else {
        R[j * k + j] = 0.0;
        for (int i = 0; i < m; ++i) Q[(size_t)i * k + j] = 0.0;
      }
    }
    for (int t = 0; t < m_nodes; ++t) {
      const long long blk = blk_of_node[members[cnt[a] + t]];
      for (int r = 0; r < b; ++r)
        for (int c = 0; c < k; ++c) p_vals[(size_t)blk * b * k + r * k + c] = Q[(size_t)(t * b + r) * k + c];
    }
    for (int r = 0; r < k; ++r)
      for (int c = 0; c < k; ++c) Bc[((size_t)a * k + r) * k + c] = R[r * k + c];
  }
  return 0;
}

}  // namespace wn
