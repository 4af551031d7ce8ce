// Host-side set-up of the element hot path: synthetic meshes, DOF numbering, the CSR
// pattern of the free DOFs, element colouring and the element -> CSR scatter map.
//
// The DOF numbering and the pattern are the reference's (generate_local_dof_ids,
// src/generate_dof_ids.cc:69-171,306-412; Psparse_ApAi, src/Psparse_ApAi.cc:238-670).  The
// numbering is the same element-order traversal; the pattern is built from the node graph
// instead of the reference's per-row lists of every element contribution followed by a
// qsort, which gives the same sorted unique columns per row with O(nnz) memory.
#include <algorithm>
#include <cstring>
#include <random>
#include <unordered_map>
#include <vector>

#include "pgf_internal.h"

namespace {

inline int64_t nodes_of(const pgf_mesh &m, int64_t e) { return m.elem_ptr[e + 1] - m.elem_ptr[e]; }

// Uniform double in [0,1) from the top 53 bits of a mt19937_64 draw; spelled out so the
// sequence does not depend on the standard library's distribution implementation.
inline double unit_draw(std::mt19937_64 &g) { return double(g() >> 11) * 0x1.0p-53; }

}  // namespace

namespace pgf {

Incidence build_incidence(const pgf_mesh &m) {
  Incidence inc;
  inc.ptr.assign(m.n_nodes + 1, 0);
  for (int64_t e = 0; e < m.n_elem; ++e)
    for (int64_t k = m.elem_ptr[e]; k < m.elem_ptr[e + 1]; ++k) inc.ptr[m.elem_conn[k] + 1]++;
  for (int64_t c = 0; c < m.n_coel; ++c)
    for (int64_t k = m.coel_ptr[c]; k < m.coel_ptr[c + 1]; ++k) inc.ptr[m.coel_conn[k] + 1]++;
  for (int64_t n = 0; n < m.n_nodes; ++n) inc.ptr[n + 1] += inc.ptr[n];
  inc.elems.resize(inc.ptr[m.n_nodes]);
  std::vector<int64_t> fill(inc.ptr.begin(), inc.ptr.end() - 1);
  for (int64_t e = 0; e < m.n_elem; ++e)
    for (int64_t k = m.elem_ptr[e]; k < m.elem_ptr[e + 1]; ++k) inc.elems[fill[m.elem_conn[k]]++] = e;
  for (int64_t c = 0; c < m.n_coel; ++c)
    for (int64_t k = m.coel_ptr[c]; k < m.coel_ptr[c + 1]; ++k)
      inc.elems[fill[m.coel_conn[k]]++] = m.n_elem + c;
  return inc;
}

int color_elements(const pgf_mesh &m, std::vector<int32_t> &color) {
  constexpr int kWords = 4;  // up to 256 colours
  std::vector<uint64_t> used(size_t(m.n_nodes) * kWords, 0);
  color.assign(m.n_elem, -1);
  int ncol = 0;
  for (int64_t e = 0; e < m.n_elem; ++e) {
    uint64_t forb[kWords] = {0, 0, 0, 0};
    for (int64_t k = m.elem_ptr[e]; k < m.elem_ptr[e + 1]; ++k)
      for (int w = 0; w < kWords; ++w) forb[w] |= used[m.elem_conn[k] * kWords + w];
    int c = -1;
    for (int w = 0; w < kWords && c < 0; ++w)
      if (~forb[w]) c = w * 64 + __builtin_ctzll(~forb[w]);
    if (c < 0) return -1;
    color[e] = c;
    ncol = std::max(ncol, c + 1);
    for (int64_t k = m.elem_ptr[e]; k < m.elem_ptr[e + 1]; ++k)
      used[m.elem_conn[k] * kWords + c / 64] |= (uint64_t(1) << (c % 64));
  }
  return ncol;
}

// Sorted free-DOF columns of every node's rows (all free DOFs of a node share one row
// pattern: the free DOFs of every element incident to the node).
static void node_columns(const pgf_mesh &m, const Incidence &inc, const int64_t *id,
                         std::vector<int64_t> &cptr, std::vector<int32_t> &cols) {
  const int64_t nn = m.n_nodes;
  std::vector<int64_t> count(nn + 1, 0);
  // pass 1: count, pass 2: fill (each node handled independently)
  auto collect = [&](int64_t n, std::vector<int32_t> &buf) {
    buf.clear();
    for (int64_t k = inc.ptr[n]; k < inc.ptr[n + 1]; ++k) {
      const int64_t e = inc.elems[k];
      const int64_t *conn;
      int64_t len;
      if (e < m.n_elem) {
        conn = m.elem_conn + m.elem_ptr[e];
        len = m.elem_ptr[e + 1] - m.elem_ptr[e];
      } else {
        const int64_t c = e - m.n_elem;
        conn = m.coel_conn + m.coel_ptr[c];
        len = m.coel_ptr[c + 1] - m.coel_ptr[c];
      }
      for (int64_t a = 0; a < len; ++a)
        for (int j = 0; j < 3; ++j) {
          const int64_t d = id[conn[a] * 3 + j];
          if (d > 0) buf.push_back(int32_t(d - 1));
        }
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t n = 0; n < nn; ++n) {
      bool has_free = id[n * 3] > 0 || id[n * 3 + 1] > 0 || id[n * 3 + 2] > 0;
      if (!has_free) continue;
      collect(n, buf);
      count[n + 1] = int64_t(buf.size());
    }
  }
  cptr.assign(nn + 1, 0);
  for (int64_t n = 0; n < nn; ++n) cptr[n + 1] = cptr[n] + count[n + 1];
  cols.resize(cptr[nn]);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t n = 0; n < nn; ++n) {
      if (cptr[n + 1] == cptr[n]) continue;
      collect(n, buf);
      std::copy(buf.begin(), buf.end(), cols.begin() + cptr[n]);
    }
  }
}

// Node-pair offsets of elements given by (ptr, conn): see build_scatter_offsets.
static int pair_offsets(int64_t n_el, const int64_t *ptr, const int64_t *conn_all, const int64_t *id,
                        const int32_t *Ap, const int32_t *Ai, std::vector<int64_t> &pair_ptr,
                        std::vector<int16_t> &off) {
  pair_ptr.assign(n_el + 1, 0);
  for (int64_t e = 0; e < n_el; ++e) {
    const int64_t nne = ptr[e + 1] - ptr[e];
    pair_ptr[e + 1] = pair_ptr[e] + nne * nne;
  }
  off.assign(pair_ptr[n_el], int16_t(-1));
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t e = 0; e < n_el; ++e) {
    const int64_t nne = ptr[e + 1] - ptr[e];
    const int64_t *conn = conn_all + ptr[e];
    for (int64_t a = 0; a < nne; ++a) {
      int64_t ra = -1;
      for (int j = 0; j < 3 && ra < 0; ++j)
        if (id[conn[a] * 3 + j] > 0) ra = id[conn[a] * 3 + j] - 1;
      if (ra < 0) continue;
      const int32_t *rb = Ai + Ap[ra], *re = Ai + Ap[ra + 1];
      for (int64_t w = 0; w < nne; ++w) {
        int64_t cw = -1;
        for (int j = 0; j < 3 && cw < 0; ++j)
          if (id[conn[w] * 3 + j] > 0) cw = id[conn[w] * 3 + j] - 1;
        if (cw < 0) continue;
        const int32_t *p = std::lower_bound(rb, re, int32_t(cw));
        if (p == re || *p != cw || (p - rb) > 32767) {
          ++bad;
          continue;
        }
        off[pair_ptr[e] + a * nne + w] = int16_t(p - rb);
      }
    }
  }
  return bad ? PGF_EINVAL : PGF_OK;
}

int build_scatter_offsets(const pgf_mesh &m, const int64_t *id, int64_t ndof,
                          const int32_t *Ap, const int32_t *Ai,
                          std::vector<int64_t> &pair_ptr, std::vector<int16_t> &off) {
  (void)ndof;
  return pair_offsets(m.n_elem, m.elem_ptr, m.elem_conn, id, Ap, Ai, pair_ptr, off);
}

int build_coel_offsets(const pgf_mesh &m, const int64_t *id, const int32_t *Ap, const int32_t *Ai,
                       std::vector<int64_t> &pair_ptr, std::vector<int16_t> &off) {
  return pair_offsets(m.n_coel, m.coel_ptr, m.coel_conn, id, Ap, Ai, pair_ptr, off);
}

int color_coels(const pgf_mesh &m, std::vector<int32_t> &color) {
  constexpr int kWords = 4;
  std::vector<uint64_t> used(size_t(m.n_nodes) * kWords, 0);
  color.assign(m.n_coel, -1);
  int ncol = 0;
  for (int64_t c = 0; c < m.n_coel; ++c) {
    uint64_t forb[kWords] = {0, 0, 0, 0};
    for (int64_t k = m.coel_ptr[c]; k < m.coel_ptr[c + 1]; ++k)
      for (int w = 0; w < kWords; ++w) forb[w] |= used[m.coel_conn[k] * kWords + w];
    int col = -1;
    for (int w = 0; w < kWords && col < 0; ++w)
      if (~forb[w]) col = w * 64 + __builtin_ctzll(~forb[w]);
    if (col < 0) return -1;
    color[c] = col;
    ncol = std::max(ncol, col + 1);
    for (int64_t k = m.coel_ptr[c]; k < m.coel_ptr[c + 1]; ++k)
      used[m.coel_conn[k] * kWords + col / 64] |= (uint64_t(1) << (col % 64));
  }
  return ncol;
}

}  // namespace pgf

extern "C" {

int pgf_mesh_box(int32_t n, int32_t etype, double jitter, uint64_t seed, int64_t *n_nodes,
                 int64_t *n_elem, int64_t *conn_len, double *X, int64_t *elem_ptr,
                 int64_t *elem_conn) {
  if (n < 1 || !(etype == PGF_TET4 || etype == PGF_TET10 || etype == PGF_HEX8))
    return PGF_EINVAL;
  const int64_t nv1 = int64_t(n) + 1;
  const int64_t nvert = nv1 * nv1 * nv1;
  const int64_t ncell = int64_t(n) * n * n;
  auto vid = [&](int64_t i, int64_t j, int64_t k) { return i + nv1 * (j + nv1 * k); };
  const int64_t epc = (etype == PGF_HEX8) ? 1 : 6;
  const int64_t nne = etype;
  const int64_t ne = ncell * epc;
  // Kuhn split: one tet per permutation of the axes, walking from corner 0 to corner 7.
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int tets[6][4];
  for (int p = 0; p < 6; ++p) {
    const int b0 = 1 << perms[p][0], b1 = 1 << perms[p][1];
    int t[4] = {0, b0, b0 | b1, 7};
    // orientation on the unit cube: det[(v1-v0),(v2-v0),(v3-v0)] must be > 0
    double v[4][3];
    for (int q = 0; q < 4; ++q)
      for (int d = 0; d < 3; ++d) v[q][d] = (t[q] >> d) & 1;
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d) {
      a[d] = v[1][d] - v[0][d];
      b[d] = v[2][d] - v[0][d];
      c[d] = v[3][d] - v[0][d];
    }
    const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                       a[2] * (b[0] * c[1] - b[1] * c[0]);
    if (det < 0) std::swap(t[1], t[2]);
    for (int q = 0; q < 4; ++q) tets[p][q] = t[q];
  }
  // vertices of cell corner c: (i + (c&1), j + (c>>1 &1), k + (c>>2 &1))
  auto corner = [&](int64_t i, int64_t j, int64_t k, int c) {
    return vid(i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1));
  };
  static const int hexc[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // FEMLIB hex node order

  // edge (mid-node) numbering for TET10: first encounter in element order
  std::unordered_map<uint64_t, int64_t> edge_id;
  int64_t nedge = 0;
  const int edge_pairs[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
  if (etype == PGF_TET10) {
    edge_id.reserve(size_t(ncell) * 8);
    for (int64_t k = 0; k < n; ++k)
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i)
          for (int p = 0; p < 6; ++p) {
            int64_t v[4];
            for (int q = 0; q < 4; ++q) v[q] = corner(i, j, k, tets[p][q]);
            for (auto &ep : edge_pairs) {
              const uint64_t lo = std::min(v[ep[0]], v[ep[1]]), hi = std::max(v[ep[0]], v[ep[1]]);
              const uint64_t key = lo * uint64_t(nvert) + hi;
              if (edge_id.emplace(key, nvert + nedge).second) ++nedge;
            }
          }
  }
  const int64_t nn_total = nvert + nedge;
  *n_nodes = nn_total;
  *n_elem = ne;
  *conn_len = ne * nne;
  if (X == nullptr) return PGF_OK;

  // vertex coordinates with jitter of interior vertices
  const double h = 1.0 / n;
  std::mt19937_64 rng(seed);
  for (int64_t k = 0; k <= n; ++k)
    for (int64_t j = 0; j <= n; ++j)
      for (int64_t i = 0; i <= n; ++i) {
        const int64_t v = vid(i, j, k);
        double x[3] = {i * h, j * h, k * h};
        const bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
        if (interior && jitter != 0.0)
          for (int d = 0; d < 3; ++d) x[d] += (2.0 * unit_draw(rng) - 1.0) * jitter * h;
        for (int d = 0; d < 3; ++d) X[v * 3 + d] = x[d];
      }
  int64_t e = 0;
  for (int64_t k = 0; k < n; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) {
        if (etype == PGF_HEX8) {
          elem_ptr[e] = e * 8;
          for (int q = 0; q < 8; ++q) elem_conn[e * 8 + q] = corner(i, j, k, hexc[q]);
          ++e;
          continue;
        }
        for (int p = 0; p < 6; ++p) {
          elem_ptr[e] = e * nne;
          int64_t v[4];
          for (int q = 0; q < 4; ++q) v[q] = elem_conn[e * nne + q] = corner(i, j, k, tets[p][q]);
          if (etype == PGF_TET10) {
            for (int q = 0; q < 6; ++q) {
              const int64_t a = v[edge_pairs[q][0]], b = v[edge_pairs[q][1]];
              const uint64_t key = uint64_t(std::min(a, b)) * uint64_t(nvert) + uint64_t(std::max(a, b));
              const int64_t m = edge_id[key];
              elem_conn[e * nne + 4 + q] = m;
              for (int d = 0; d < 3; ++d) X[m * 3 + d] = 0.5 * (X[a * 3 + d] + X[b * 3 + d]);
            }
          }
          ++e;
        }
      }
  elem_ptr[ne] = ne * nne;
  return PGF_OK;
}

int pgf_generate_dof_ids(const pgf_mesh *mesh, int64_t *id, int64_t *ndof) {
  if (!mesh || !id || !ndof) return PGF_EINVAL;
  const pgf_mesh &m = *mesh;
  std::vector<uint8_t> visited(size_t(m.n_nodes) * 3, 0);
  int64_t next = 1;
  auto visit_nodes = [&](const int64_t *conn, int64_t len) {
    for (int64_t a = 0; a < len; ++a) {
      const int64_t node = conn[a];
      for (int j = 0; j < 3; ++j) {
        const int64_t k = node * 3 + j;
        if (visited[k]) continue;
        visited[k] = 1;
        if (id[k] < 0) continue;          // prescribed: keep -index
        if (id[k] == 1) { id[k] = 0; continue; }  // fixed
        id[k] = next++;                   // free
      }
    }
  };
  for (int64_t e = 0; e < m.n_elem; ++e)
    visit_nodes(m.elem_conn + m.elem_ptr[e], m.elem_ptr[e + 1] - m.elem_ptr[e]);
  for (int64_t c = 0; c < m.n_coel; ++c)
    visit_nodes(m.coel_conn + m.coel_ptr[c], m.coel_ptr[c + 1] - m.coel_ptr[c]);
  for (size_t k = 0; k < visited.size(); ++k)
    if (!visited[k]) return PGF_EINVAL;  // hanging node (no periodic nodes here)
  *ndof = next - 1;
  return PGF_OK;
}

int pgf_psparse_apai(const pgf_mesh *mesh, const int64_t *id, int64_t ndof, int32_t *Ap,
                     int32_t *Ai, int64_t *nnz) {
  if (!mesh || !id || !Ap || !nnz) return PGF_EINVAL;
  const pgf_mesh &m = *mesh;
  pgf::Incidence inc = pgf::build_incidence(m);
  std::vector<int64_t> cptr;
  std::vector<int32_t> cols;
  pgf::node_columns(m, inc, id, cptr, cols);
  // row r (free DOF id r+1) belongs to node r_node[r]
  std::vector<int64_t> row_node(ndof, -1);
  for (int64_t n = 0; n < m.n_nodes; ++n)
    for (int j = 0; j < 3; ++j) {
      const int64_t d = id[n * 3 + j];
      if (d > 0) {
        if (d > ndof) return PGF_EINVAL;
        row_node[d - 1] = n;
      }
    }
  int64_t total = 0;
  Ap[0] = 0;
  for (int64_t r = 0; r < ndof; ++r) {
    if (row_node[r] < 0) return PGF_EINVAL;
    total += cptr[row_node[r] + 1] - cptr[row_node[r]];
    if (total > INT32_MAX) return PGF_EINVAL;
    Ap[r + 1] = int32_t(total);
  }
  *nnz = total;
  if (Ai == nullptr) return PGF_OK;
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < ndof; ++r) {
    const int64_t n = row_node[r];
    std::copy(cols.begin() + cptr[n], cols.begin() + cptr[n + 1], Ai + Ap[r]);
  }
  return PGF_OK;
}

}  // extern "C"

