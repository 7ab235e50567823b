// Host-side setup of the solve path (native C++, runs once per mesh):
//   * reverse Cuthill-McKee ordering of a node graph,
//   * the node graph / 3x3 block pattern of a C3D10 mesh with the Dirichlet
//     pattern rule,
//   * the per-block contribution lists that make GPU assembly a deterministic
//     gather, and the node -> (element, local node) incidence lists;
//   * the block-sparse products of the smoothed-aggregation set-up (A P and
//     R (A P)), the block transpose, the in-place symmetrisation of a
//     Galerkin operator and the block SpMV of its eigenvalue estimate, on
//     host threads (one contiguous range of block rows each; every output
//     entry is summed in a fixed order, so results do not depend on the
//     thread count);
//   * the patch layout of the matrix-free C3D10 operator (elements in Morton
//     order, cut into patches; local connectivity, per-node slot lists and
//     each node's partial sums).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <utility>
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

// ---------------------------------------------------------------- threaded block-sparse set-up kernels
// Host threads to use when the caller asks for 0: OMP_NUM_THREADS when set
// (a job's CPU share), else one per hardware thread.
static int default_threads() {
  if (const char* e = std::getenv("OMP_NUM_THREADS")) {
    const int v = std::atoi(e);
    if (v > 0) return v;
  }
  return (int)std::thread::hardware_concurrency();
}

// Runs fn(row_begin, row_end) on `threads` host threads over [0, n) (0 = the
// default above, at most 64).
template <class F>
static void parallel_rows(int n, int threads, F fn) {
  int t = threads > 0 ? threads : default_threads();
  t = std::max(1, std::min({t, 64, std::max(1, n / 64)}));
  if (t == 1) {
    fn(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const int per = (n + t - 1) / t;
  for (int k = 0; k < t; ++k) {
    const int a = k * per, b = std::min(n, a + per);
    if (a < b) pool.emplace_back(fn, a, b);
  }
  for (auto& th : pool) th.join();
}

// C = A B for block-CSR matrices (row-major blocks): A has nr block rows of
// ra x ca blocks, B has ca x cb blocks and nc block columns.  With c_idx ==
// nullptr only the structure pass runs: c_ptr[0..nr] receives the prefix sums
// of the block counts of C's rows, and the total is returned.  With c_idx
// and c_val (c_ptr[nr] entries) the products are accumulated: each row of C
// in a dense accumulator over B's block columns, its columns sorted, each
// block summed over A's row in storage order.
long long bsr_spgemm(int nr, int ra, int ca, int cb, int nc, const int* a_ptr, const int* a_idx, const double* a_val,
                     const int* b_ptr, const int* b_idx, const double* b_val, int* c_ptr, int* c_idx, double* c_val,
                     int threads) {
  const int bbA = ra * ca, bbB = ca * cb, bbC = ra * cb;
  if (!c_idx) {
    std::vector<int> cnt(nr, 0);
    parallel_rows(nr, threads, [&](int r0, int r1) {
      std::vector<int> mark(nc, -1);
      for (int i = r0; i < r1; ++i) {
        int c = 0;
        for (int ka = a_ptr[i]; ka < a_ptr[i + 1]; ++ka) {
          const int j = a_idx[ka];
          for (int kb = b_ptr[j]; kb < b_ptr[j + 1]; ++kb) {
            const int col = b_idx[kb];
            if (mark[col] != i) {
              mark[col] = i;
              ++c;
            }
          }
        }
        cnt[i] = c;
      }
    });
    long long tot = 0;
    c_ptr[0] = 0;
    for (int i = 0; i < nr; ++i) {
      tot += cnt[i];
      c_ptr[i + 1] = (int)tot;
    }
    return tot;
  }
  parallel_rows(nr, threads, [&](int r0, int r1) {
    std::vector<int> pos(nc, -1);
    std::vector<int> cols;
    std::vector<double> acc;
    for (int i = r0; i < r1; ++i) {
      cols.clear();
      for (int ka = a_ptr[i]; ka < a_ptr[i + 1]; ++ka) {
        const int j = a_idx[ka];
        for (int kb = b_ptr[j]; kb < b_ptr[j + 1]; ++kb) {
          const int col = b_idx[kb];
          if (pos[col] < 0) {
            pos[col] = 0;
            cols.push_back(col);
          }
        }
      }
      std::sort(cols.begin(), cols.end());
      for (size_t q = 0; q < cols.size(); ++q) pos[cols[q]] = (int)q;
      acc.assign(cols.size() * bbC, 0.0);
      for (int ka = a_ptr[i]; ka < a_ptr[i + 1]; ++ka) {
        const int j = a_idx[ka];
        const double* Ab = a_val + (size_t)ka * bbA;
        for (int kb = b_ptr[j]; kb < b_ptr[j + 1]; ++kb) {
          const double* Bb = b_val + (size_t)kb * bbB;
          double* Cb = acc.data() + (size_t)pos[b_idx[kb]] * bbC;
          for (int r = 0; r < ra; ++r)
            for (int m = 0; m < ca; ++m) {
              const double av = Ab[r * ca + m];
              if (av == 0.0) continue;
              for (int c = 0; c < cb; ++c) Cb[r * cb + c] += av * Bb[m * cb + c];
            }
        }
      }
      const int base = c_ptr[i];
      for (size_t q = 0; q < cols.size(); ++q) {
        c_idx[base + q] = cols[q];
        std::memcpy(c_val + (size_t)(base + q) * bbC, acc.data() + q * bbC, sizeof(double) * bbC);
        pos[cols[q]] = -1;
      }
    }
  });
  return c_ptr[nr];
}

// T = A^T for a block-CSR A (nr block rows, nc block columns, ra x ca blocks):
// t_ptr[nc + 1], t_idx / t_val[nnzb] (rows of T sorted by column, blocks
// transposed).
void bsr_transpose(int nr, int nc, int ra, int ca, const int* a_ptr, const int* a_idx, const double* a_val,
                   int* t_ptr, int* t_idx, double* t_val) {
  const long long nnzb = a_ptr[nr];
  std::vector<int> cnt(nc + 1, 0);
  for (long long k = 0; k < nnzb; ++k) cnt[a_idx[k] + 1]++;
  for (int c = 0; c < nc; ++c) cnt[c + 1] += cnt[c];
  std::memcpy(t_ptr, cnt.data(), sizeof(int) * (nc + 1));
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int i = 0; i < nr; ++i)
    for (int k = a_ptr[i]; k < a_ptr[i + 1]; ++k) {
      const int c = a_idx[k], dst = fill[c]++;
      t_idx[dst] = i;
      const double* src = a_val + (size_t)k * ra * ca;
      double* out = t_val + (size_t)dst * ra * ca;
      for (int r = 0; r < ra; ++r)
        for (int q = 0; q < ca; ++q) out[q * ra + r] = src[r * ca + q];
    }
}

// In place A <- (A + A^T) / 2 for a square block-CSR matrix with a
// structurally symmetric pattern and sorted columns; each pair of mirror
// blocks is averaged once.  Zero diagonal entries are then set to one (a
// coarse DOF that no fine DOF interpolates).  Returns the number of entries
// so set, or -1 when a mirror block is missing.
long long bsr_symmetrize(int n, int b, const int* ptr, const int* idx, double* val, int threads) {
  const int bb = b * b;
  std::vector<long long> fixed(std::max(1, n), 0);
  std::vector<char> bad(std::max(1, n), 0);
  parallel_rows(n, threads, [&](int r0, int r1) {
    for (int i = r0; i < r1; ++i) {
      for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int j = idx[k];
        if (j < i) continue;
        double* A = val + (size_t)k * bb;
        if (j == i) {
          for (int r = 0; r < b; ++r)
            for (int c = r + 1; c < b; ++c) {
              const double m = 0.5 * (A[r * b + c] + A[c * b + r]);
              A[r * b + c] = m;
              A[c * b + r] = m;
            }
          for (int r = 0; r < b; ++r)
            if (A[r * b + r] == 0.0) {
              A[r * b + r] = 1.0;
              fixed[i]++;
            }
          continue;
        }
        const int* lo = idx + ptr[j];
        const int* hi = idx + ptr[j + 1];
        const int* at = std::lower_bound(lo, hi, i);
        if (at == hi || *at != i) {
          bad[i] = 1;
          continue;
        }
        double* T = val + (size_t)(ptr[j] + (at - lo)) * bb;
        // block (i, j) and block (j, i) = its transpose: average them
        for (int r = 0; r < b; ++r)
          for (int c = 0; c < b; ++c) {
            const double m = 0.5 * (A[r * b + c] + T[c * b + r]);
            A[r * b + c] = m;
            T[c * b + r] = m;
          }
      }
    }
  });
  long long tot = 0;
  for (int i = 0; i < n; ++i) {
    if (bad[i]) return -1;
    tot += fixed[i];
  }
  return tot;
}

// y = A x for a square block-CSR matrix (b x b blocks) on host threads.
void bsr_spmv_host(int n, int b, const int* ptr, const int* idx, const double* val, const double* x, double* y,
                   int threads) {
  const int bb = b * b;
  parallel_rows(n, threads, [&](int r0, int r1) {
    for (int i = r0; i < r1; ++i) {
      double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        const double* A = val + (size_t)k * bb;
        const double* xv = x + (size_t)idx[k] * b;
        for (int r = 0; r < b; ++r)
          for (int c = 0; c < b; ++c) acc[r] += A[r * b + c] * xv[c];
      }
      for (int r = 0; r < b; ++r) y[(size_t)i * b + r] = acc[r];
    }
  });
}

// ---------------------------------------------------------------- matrix-free patch layout
// Morton order of C3D10 elements: the elements grouped by connected
// component (elements that share a corner, found by union-find; components
// in the order of their first element), and inside a component sorted by the
// Morton code of the centroid of their four corners, quantised to 21 bits
// per axis on the component's bounding cube (bits interleaved, x lowest;
// ties in element order).  Grouping first keeps separate bodies that overlap
// in space (a batch of meshes in one system) out of each other's patches.
static unsigned long long spread3(unsigned long long v) {  // bit i -> bit 3 i (21 bits)
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

static int uf_find(std::vector<int>& par, int a) {
  while (par[(size_t)a] != a) {
    par[(size_t)a] = par[(size_t)par[(size_t)a]];
    a = par[(size_t)a];
  }
  return a;
}

void morton_order(int M, const int* conn, const double* X, int* order) {
  // components over the corner nodes
  int nmax = 0;
  for (long long q = 0; q < 10LL * M; ++q) nmax = std::max(nmax, conn[q] + 1);
  std::vector<int> par((size_t)nmax);
  for (int i = 0; i < nmax; ++i) par[(size_t)i] = i;
  for (int e = 0; e < M; ++e) {
    const int r0 = uf_find(par, conn[(size_t)e * 10]);
    for (int i = 1; i < 4; ++i) {
      const int ri = uf_find(par, conn[(size_t)e * 10 + i]);
      if (ri != r0) par[(size_t)std::max(ri, r0)] = std::min(ri, r0);
    }
  }
  // component index of each element, components numbered by first element
  std::vector<int> comp((size_t)M), label((size_t)nmax, -1);
  int nc = 0;
  for (int e = 0; e < M; ++e) {
    const int r = uf_find(par, conn[(size_t)e * 10]);
    if (label[(size_t)r] < 0) label[(size_t)r] = nc++;
    comp[(size_t)e] = label[(size_t)r];
  }
  std::vector<double> c((size_t)M * 3);
  std::vector<double> lo((size_t)nc * 3, 1e300), hi((size_t)nc * 3, -1e300);
  for (int e = 0; e < M; ++e)
    for (int a = 0; a < 3; ++a) {
      double v = 0.0;
      for (int i = 0; i < 4; ++i) v += X[(size_t)conn[(size_t)e * 10 + i] * 3 + a];
      v *= 0.25;
      c[(size_t)e * 3 + a] = v;
      const size_t k = (size_t)comp[(size_t)e] * 3 + a;
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  const double top = double((1 << 21) - 1);
  std::vector<double> q((size_t)nc);
  for (int k = 0; k < nc; ++k) {
    double ext = 0.0;
    for (int a = 0; a < 3; ++a) ext = std::max(ext, hi[(size_t)k * 3 + a] - lo[(size_t)k * 3 + a]);
    q[(size_t)k] = top / (ext > 0.0 ? ext : 1.0);
  }
  struct Key {
    int comp;
    unsigned long long code;
    int e;
    bool operator<(const Key& o) const {
      return comp != o.comp ? comp < o.comp : (code != o.code ? code < o.code : e < o.e);
    }
  };
  std::vector<Key> key((size_t)M);
  for (int e = 0; e < M; ++e) {
    const int k = comp[(size_t)e];
    unsigned long long code = 0;
    for (int a = 0; a < 3; ++a) {
      const double t = std::min(std::max((c[(size_t)e * 3 + a] - lo[(size_t)k * 3 + a]) * q[(size_t)k], 0.0), top);
      code |= spread3((unsigned long long)t) << a;
    }
    key[(size_t)e] = {k, code, e};
  }
  std::sort(key.begin(), key.end());
  for (int e = 0; e < M; ++e) order[e] = key[(size_t)e].e;
}

// Patch layout: slot s holds element order[s]; patch p the slots
// [p P, min(M, (p + 1) P)).  A patch's local nodes are the distinct global
// nodes of its elements, ascending.  mf_patch_sizes returns the number of
// local nodes of all patches and the most of one patch (lmax).
// mf_patches fills, for a per-patch stride >= lmax + 1:
//   pnode[p stride + j]  global node of local node j (-1 past the last one);
//   pconn[10 s + i]      local node of corner i of slot s;
//   pslot[10 p P ...]    the local nodes' slot codes 10 (s - p P) + i, each
//                        node's list ascending, local node j's from
//                        poff[p stride + j] to poff[p stride + j + 1] (the
//                        entries past the last node hold 10 ne);
//   nhead[4 n]           node n's partial slots p stride + j in patch order:
//                        the first three (-1 where absent) and their count,
//                        or, with more than three, -(k + 1) where
//                        novf[k] = count - 3 and novf[k + 1 ...] the rest.
// Returns the length of novf, or -1 on a corner outside [0, N), a repeated
// element, P outside [1, 6553] or stride <= lmax.
static bool mf_patch_check(int M, int N, const int* order, int P) {
  if (M < 0 || N < 0 || P < 1 || P > 6553) return false;
  std::vector<char> seen((size_t)M, 0);
  for (int sl = 0; sl < M; ++sl) {
    const int e = order[sl];
    if (e < 0 || e >= M || seen[(size_t)e]) return false;
    seen[(size_t)e] = 1;
  }
  return true;
}

// local nodes of patch p (ascending) into ln; false on a corner outside [0, N)
static bool patch_nodes(int M, int N, const int* conn, const int* order, int P, int p, std::vector<int>& ln) {
  const int s0 = p * P, s1 = std::min(M, s0 + P);
  ln.clear();
  for (int sl = s0; sl < s1; ++sl)
    for (int i = 0; i < 10; ++i) {
      const int g = conn[(size_t)order[sl] * 10 + i];
      if (g < 0 || g >= N) return false;
      ln.push_back(g);
    }
  std::sort(ln.begin(), ln.end());
  ln.erase(std::unique(ln.begin(), ln.end()), ln.end());
  return true;
}

long long mf_patch_sizes(int M, int N, const int* conn, const int* order, int P, int* lmax) {
  if (!mf_patch_check(M, N, order, P)) return -1;
  const int n_patch = (M + P - 1) / P;
  std::vector<int> ln;
  long long nl = 0;
  int lm = 0;
  for (int p = 0; p < n_patch; ++p) {
    if (!patch_nodes(M, N, conn, order, P, p, ln)) return -1;
    nl += (long long)ln.size();
    lm = std::max(lm, (int)ln.size());
  }
  *lmax = lm;
  return nl;
}

long long mf_patches(int M, int N, const int* conn, const int* order, int P, int stride, int* pnode,
                     unsigned short* pconn, unsigned short* poff, unsigned short* pslot, int* nhead, int* novf) {
  if (!mf_patch_check(M, N, order, P) || stride < 2) return -1;
  const int n_patch = (M + P - 1) / P;
  std::vector<int> loc((size_t)N, -1), ln, cnt;
  std::vector<std::vector<int>> parts;  // per node, only while building
  std::vector<int> count((size_t)N, 0);
  for (int p = 0; p < n_patch; ++p) {
    const int s0 = p * P, s1 = std::min(M, s0 + P), ne = s1 - s0;
    if (!patch_nodes(M, N, conn, order, P, p, ln)) return -1;
    const int L = (int)ln.size();
    if (L >= stride) return -1;
    int* pn = pnode + (size_t)p * stride;
    unsigned short* po = poff + (size_t)p * stride;
    for (int j = 0; j < L; ++j) {
      loc[(size_t)ln[(size_t)j]] = j;
      pn[j] = ln[(size_t)j];
    }
    for (int j = L; j < stride; ++j) {
      pn[j] = -1;
      po[j] = (unsigned short)(10 * ne);
    }
    for (int sl = s0; sl < s1; ++sl)
      for (int i = 0; i < 10; ++i)
        pconn[(size_t)sl * 10 + i] = (unsigned short)loc[(size_t)conn[(size_t)order[sl] * 10 + i]];
    // counting sort of the slot codes by local node: each list ascending
    cnt.assign((size_t)L + 1, 0);
    const unsigned short* pc = pconn + (size_t)s0 * 10;
    for (int q = 0; q < 10 * ne; ++q) ++cnt[(size_t)pc[q] + 1];
    for (int j = 0; j < L; ++j) cnt[(size_t)j + 1] += cnt[(size_t)j];
    for (int j = 0; j < L; ++j) po[j] = (unsigned short)cnt[(size_t)j];
    for (int q = 0; q < 10 * ne; ++q) pslot[(size_t)s0 * 10 + cnt[(size_t)pc[q]]++] = (unsigned short)q;
    for (int j = 0; j < L; ++j) {
      const int g = ln[(size_t)j];
      loc[(size_t)g] = -1;
      const int slot = p * stride + j;
      int* h = nhead + (size_t)g * 4;
      if (count[(size_t)g] < 3) {
        h[count[(size_t)g]] = slot;
      } else {
        if (parts.empty()) parts.resize((size_t)N);
        parts[(size_t)g].push_back(slot);
      }
      ++count[(size_t)g];
    }
  }
  long long nov = 0;
  for (int n = 0; n < N; ++n) {
    int* h = nhead + (size_t)n * 4;
    const int c = count[(size_t)n];
    for (int k = c; k < 3; ++k) h[k] = -1;
    if (c <= 3) {
      h[3] = c;
    } else {
      h[3] = -(int)(nov + 1);
      novf[nov++] = c - 3;
      for (int v : parts[(size_t)n]) novf[nov++] = v;
    }
  }
  return nov;
}

}  // namespace wn
