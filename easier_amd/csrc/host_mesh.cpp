// Host-side mesh generation and index renumbering.
//
// esr_mesh_tri4 produces the same arrays as the reference generator
// (csrc/triangular_mesh.cpp:10-142): 4 triangles per square around the square
// centre, cells 4q..4q+3 (bottom, right, top, left) of square q = j*n + i,
// 8 intra-square directed edges per square, then the four families of
// inter-square edges (down, up, left, right) at the offsets of :85-130, and the
// boundary faces in the order bottom, top, left, right. Where several squares
// write one shared corner the reference's result depends on write order; this
// implementation fixes it to the sequential (single-thread) outcome.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <numeric>
#include <vector>

#include "esr_host.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform in [0, 1) with 53 random bits
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

} // namespace

ESR_HAPI int esr_mesh_sizes(int kind, int64_t n, int64_t *np_, int64_t *nc, int64_t *ne,
                            int64_t *nbc) {
  ESR_HREQUIRE(n >= 1, "esr_mesh_sizes: n must be >= 1");
  ESR_HREQUIRE(np_ && nc && ne && nbc, "esr_mesh_sizes: null pointer");
  if (kind == 4) {
    *np_ = (n + 1) * (n + 1) + n * n;
    *nc = 4 * n * n;
    *ne = 12 * n * n - 4 * n;
    *nbc = 4 * n;
  } else if (kind == 2) {
    *np_ = (n + 1) * (n + 1);
    *nc = 2 * n * n;
    *ne = 6 * n * n - 4 * n;
    *nbc = 4 * n;
  } else {
    esr::set_error("esr_mesh_sizes: kind must be 2 or 4, got %d", kind);
    return ESR_ERR_INVALID;
  }
  return ESR_OK;
}

ESR_HAPI int esr_mesh_tri4(int64_t n, int64_t *src, int64_t *dst, int64_t *cells,
                           int64_t *bcells, int64_t *bpoints, double *points) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(n >= 1, "esr_mesh_tri4: n must be >= 1");
  ESR_HREQUIRE(src && dst && cells && bcells && bpoints && points,
               "esr_mesh_tri4: null pointer");
  const double d = 1.0 / (double)n;
  const int64_t W = n + 1;                  // grid points per row
  const int64_t centre0 = W * W;            // first centre point
  const int64_t e_down = 8 * n * n;         // offsets of the four families
  const int64_t e_up = e_down + n * (n - 1);
  const int64_t e_left = e_up + n * (n - 1);
  const int64_t e_right = e_left + n * (n - 1);

  // Grid points: the coordinate written last in sequential square order.
  for (int64_t J = 0; J <= n; ++J) {
    for (int64_t I = 0; I <= n; ++I) {
      const double x = I < n ? (double)I * d : (double)(I - 1) * d + d;
      const double y = J < n ? (double)J * d : (double)(J - 1) * d + d;
      points[2 * (J * W + I)] = x;
      points[2 * (J * W + I) + 1] = y;
    }
  }
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < n; ++i) {
      const int64_t q = j * n + i, c = 4 * q;
      const int64_t p00 = j * W + i, p01 = (j + 1) * W + i;
      const int64_t p11 = (j + 1) * W + i + 1, p10 = j * W + i + 1;
      const int64_t pc = centre0 + q;
      points[2 * pc] = (double)i * d + d / 2;
      points[2 * pc + 1] = (double)j * d + d / 2;

      const int64_t tri[4][3] = {{p00, p10, pc}, {p11, pc, p10}, {p11, p01, pc}, {p00, pc, p01}};
      for (int k = 0; k < 4; ++k)
        for (int v = 0; v < 3; ++v) cells[3 * (c + k) + v] = tri[k][v];

      // intra-square ring k -> k+1 (mod 4), both directions
      for (int k = 0; k < 4; ++k) {
        const int64_t a = c + k, b = c + (k + 1) % 4;
        src[8 * q + 2 * k] = a;
        dst[8 * q + 2 * k] = b;
        src[8 * q + 2 * k + 1] = b;
        dst[8 * q + 2 * k + 1] = a;
      }
      if (j > 0) { // bottom triangle -> top triangle of the square below
        src[e_down + n * (j - 1) + i] = c;
        dst[e_down + n * (j - 1) + i] = c - 4 * n + 2;
      } else {
        bcells[i] = c;
        bpoints[2 * i] = p00;
        bpoints[2 * i + 1] = p10;
      }
      if (j < n - 1) { // top triangle -> bottom triangle of the square above
        src[e_up + n * j + i] = c + 2;
        dst[e_up + n * j + i] = c + 4 * n;
      } else {
        bcells[n + i] = c + 2;
        bpoints[2 * (n + i)] = p01;
        bpoints[2 * (n + i) + 1] = p11;
      }
      if (i > 0) { // left triangle -> right triangle of the square to the left
        src[e_left + (n - 1) * j + i - 1] = c + 3;
        dst[e_left + (n - 1) * j + i - 1] = c - 3;
      } else {
        bcells[2 * n + j] = c + 3;
        bpoints[2 * (2 * n + j)] = p00;
        bpoints[2 * (2 * n + j) + 1] = p01;
      }
      if (i < n - 1) { // right triangle -> left triangle of the square to the right
        src[e_right + (n - 1) * j + i] = c + 1;
        dst[e_right + (n - 1) * j + i] = c + 7;
      } else {
        bcells[3 * n + j] = c + 1;
        bpoints[2 * (3 * n + j)] = p11;
        bpoints[2 * (3 * n + j) + 1] = p10;
      }
    }
  }
  return ESR_OK;
  ESR_HGUARD_END
}

// 2 triangles per square: cell 2q = (p00, p10, p11) (below the diagonal),
// cell 2q+1 = (p00, p11, p01) (above it); both counter-clockwise.
// Directed dual edges: 2q<->2q+1 inside each square (2n^2), then the vertical
// faces between squares (i,j)|(i+1,j) (2n(n-1)), then the horizontal faces
// between (i,j) and (i,j+1) (2n(n-1)). Boundary faces: bottom, top, left,
// right, each in increasing i or j.
ESR_HAPI int esr_mesh_tri2(int64_t n, double jitter, uint64_t seed, int64_t *src,
                           int64_t *dst, int64_t *cells, int64_t *bcells, int64_t *bpoints,
                           double *points) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(n >= 1, "esr_mesh_tri2: n must be >= 1");
  ESR_HREQUIRE(jitter >= 0.0 && jitter < 0.5, "esr_mesh_tri2: jitter must be in [0, 0.5)");
  ESR_HREQUIRE(src && dst && cells && bcells && bpoints && points,
               "esr_mesh_tri2: null pointer");
  const int64_t W = n + 1;
  const double h = 1.0 / (double)n;
  SplitMix64 rng{seed};
  for (int64_t J = 0; J <= n; ++J) {
    for (int64_t I = 0; I <= n; ++I) {
      double x = (double)I / (double)n, y = (double)J / (double)n;
      if (jitter > 0.0 && I > 0 && I < n && J > 0 && J < n) {
        x += (2.0 * rng.uniform() - 1.0) * jitter * h;
        y += (2.0 * rng.uniform() - 1.0) * jitter * h;
      }
      points[2 * (J * W + I)] = x;
      points[2 * (J * W + I) + 1] = y;
    }
  }
  const int64_t e_vert = 2 * n * n;
  const int64_t e_horz = e_vert + 2 * n * (n - 1);
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < n; ++i) {
      const int64_t q = j * n + i, c = 2 * q;
      const int64_t p00 = j * W + i, p10 = j * W + i + 1;
      const int64_t p01 = (j + 1) * W + i, p11 = (j + 1) * W + i + 1;
      cells[3 * c] = p00;
      cells[3 * c + 1] = p10;
      cells[3 * c + 2] = p11;
      cells[3 * (c + 1)] = p00;
      cells[3 * (c + 1) + 1] = p11;
      cells[3 * (c + 1) + 2] = p01;
      src[2 * q] = c;
      dst[2 * q] = c + 1;
      src[2 * q + 1] = c + 1;
      dst[2 * q + 1] = c;
      if (i < n - 1) { // face x = i+1: lower triangle here, upper triangle to the right
        const int64_t k = e_vert + 2 * ((n - 1) * j + i);
        src[k] = c;
        dst[k] = c + 3;
        src[k + 1] = c + 3;
        dst[k + 1] = c;
      }
      if (j < n - 1) { // face y = j+1: upper triangle here, lower triangle above
        const int64_t k = e_horz + 2 * (n * j + i);
        src[k] = c + 1;
        dst[k] = c + 2 * n;
        src[k + 1] = c + 2 * n;
        dst[k + 1] = c + 1;
      }
      if (j == 0) {
        bcells[i] = c;
        bpoints[2 * i] = p00;
        bpoints[2 * i + 1] = p10;
      }
      if (j == n - 1) {
        bcells[n + i] = c + 1;
        bpoints[2 * (n + i)] = p01;
        bpoints[2 * (n + i) + 1] = p11;
      }
      if (i == 0) {
        bcells[2 * n + j] = c + 1;
        bpoints[2 * (2 * n + j)] = p00;
        bpoints[2 * (2 * n + j) + 1] = p01;
      }
      if (i == n - 1) {
        bcells[3 * n + j] = c;
        bpoints[2 * (3 * n + j)] = p10;
        bpoints[2 * (3 * n + j) + 1] = p11;
      }
    }
  }
  return ESR_OK;
  ESR_HGUARD_END
}

// ---------------------------------------------------------------------------
// renumbering
// ---------------------------------------------------------------------------
ESR_HAPI int esr_csr_from_idx(const int64_t *idx, int64_t m, int64_t n, int64_t *rowptr,
                              int64_t *perm) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(m >= 0 && n >= 0, "esr_csr_from_idx: negative size");
  ESR_HREQUIRE(rowptr != nullptr && (m == 0 || (idx && perm)),
               "esr_csr_from_idx: null pointer");
  std::fill(rowptr, rowptr + n + 1, 0);
  for (int64_t e = 0; e < m; ++e) {
    const int64_t r = idx[e];
    if (r < 0 || r >= n) {
      esr::set_error("esr_csr_from_idx: idx[%lld] = %lld outside [0, %lld)", (long long)e,
                     (long long)r, (long long)n);
      return ESR_ERR_INVALID;
    }
    rowptr[r + 1] += 1;
  }
  for (int64_t r = 0; r < n; ++r) rowptr[r + 1] += rowptr[r];
  std::vector<int64_t> cursor(rowptr, rowptr + n);
  for (int64_t e = 0; e < m; ++e) perm[cursor[idx[e]]++] = e;
  return ESR_OK;
  ESR_HGUARD_END
}

namespace {

struct Graph {
  std::vector<int64_t> ptr, adj;
};

Graph build_sym(int64_t n, const int64_t *a, const int64_t *b, int64_t m) {
  std::vector<std::pair<int64_t, int64_t>> pairs;
  pairs.reserve(2 * (size_t)m);
  for (int64_t k = 0; k < m; ++k) {
    if (a[k] == b[k]) continue;
    pairs.emplace_back(a[k], b[k]);
    pairs.emplace_back(b[k], a[k]);
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  Graph g;
  g.ptr.assign(n + 1, 0);
  for (auto &p : pairs) g.ptr[p.first + 1]++;
  for (int64_t v = 0; v < n; ++v) g.ptr[v + 1] += g.ptr[v];
  g.adj.resize(pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) g.adj[k] = pairs[k].second;
  return g;
}

// BFS from s over unvisited-in-`comp_mark` vertices; returns level structure
// depth and the last level's vertices.
int64_t bfs_levels(const Graph &g, int64_t s, std::vector<int64_t> &level,
                   std::vector<int64_t> &last) {
  std::deque<int64_t> q;
  std::vector<int64_t> touched;
  level[s] = 0;
  q.push_back(s);
  touched.push_back(s);
  int64_t depth = 0;
  while (!q.empty()) {
    const int64_t v = q.front();
    q.pop_front();
    depth = std::max(depth, level[v]);
    for (int64_t k = g.ptr[v]; k < g.ptr[v + 1]; ++k) {
      const int64_t w = g.adj[k];
      if (level[w] < 0) {
        level[w] = level[v] + 1;
        q.push_back(w);
        touched.push_back(w);
      }
    }
  }
  last.clear();
  for (int64_t v : touched)
    if (level[v] == depth) last.push_back(v);
  for (int64_t v : touched) level[v] = -1;
  return depth;
}

} // namespace

ESR_HAPI int esr_rcm(int64_t n, const int64_t *a, const int64_t *b, int64_t m,
                     int64_t *order) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(n >= 0 && m >= 0, "esr_rcm: negative size");
  ESR_HREQUIRE(order != nullptr && (m == 0 || (a && b)), "esr_rcm: null pointer");
  for (int64_t k = 0; k < m; ++k) {
    if (a[k] < 0 || a[k] >= n || b[k] < 0 || b[k] >= n) {
      esr::set_error("esr_rcm: pair %lld = (%lld, %lld) outside [0, %lld)", (long long)k,
                     (long long)a[k], (long long)b[k], (long long)n);
      return ESR_ERR_INVALID;
    }
  }
  const Graph g = build_sym(n, a, b, m);
  auto deg = [&](int64_t v) { return g.ptr[v + 1] - g.ptr[v]; };
  std::vector<char> done(n, 0);
  std::vector<int64_t> level(n, -1), last;
  std::vector<int64_t> cm;
  cm.reserve(n);
  std::vector<int64_t> nb;
  for (int64_t root = 0; root < n; ++root) {
    if (done[root]) continue;
    // pseudo-peripheral start: repeat BFS from a minimum-degree vertex of the
    // last level while the eccentricity grows (George & Liu).
    int64_t s = root;
    int64_t depth = bfs_levels(g, s, level, last);
    for (int it = 0; it < 8; ++it) {
      int64_t cand = last[0];
      for (int64_t v : last)
        if (deg(v) < deg(cand) || (deg(v) == deg(cand) && v < cand)) cand = v;
      const int64_t d2 = bfs_levels(g, cand, level, last);
      if (d2 <= depth) break;
      depth = d2;
      s = cand;
    }
    // Cuthill-McKee BFS, neighbours by increasing degree (ties by id)
    const size_t head0 = cm.size();
    cm.push_back(s);
    done[s] = 1;
    for (size_t h = head0; h < cm.size(); ++h) {
      const int64_t v = cm[h];
      nb.clear();
      for (int64_t k = g.ptr[v]; k < g.ptr[v + 1]; ++k)
        if (!done[g.adj[k]]) nb.push_back(g.adj[k]);
      std::sort(nb.begin(), nb.end(), [&](int64_t x, int64_t y) {
        return deg(x) != deg(y) ? deg(x) < deg(y) : x < y;
      });
      for (int64_t w : nb) {
        done[w] = 1;
        cm.push_back(w);
      }
    }
  }
  for (int64_t k = 0; k < n; ++k) order[k] = cm[n - 1 - k];
  return ESR_OK;
  ESR_HGUARD_END
}

ESR_HAPI int esr_block_order(int64_t n, const int64_t *a, const int64_t *b, int64_t m,
                             int64_t block, int64_t *order) {
  ESR_HGUARD_BEGIN
  ESR_HREQUIRE(block >= 1, "esr_block_order: block size %lld < 1", (long long)block);
  // the RCM sequence decides where blocks start and in which order they come
  std::vector<int64_t> rcm(n);
  const int rc = esr_rcm(n, a, b, m, n ? rcm.data() : order);
  if (rc != ESR_OK) return rc;
  const Graph g = build_sym(n, a, b, m);
  std::vector<char> taken(n, 0);
  int64_t out = 0;
  std::vector<int64_t> queue;
  queue.reserve(block);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t s = rcm[k];
    if (taken[s]) continue;
    // a breadth-first ball of untaken vertices around s, in BFS order
    // (neighbours in adjacency order), as large as the rest of the current
    // run of `block` vertices: balls that end early are followed by balls
    // that fill the run, so every run starts a ball
    const int64_t limit = block - out % block;
    queue.clear();
    queue.push_back(s);
    taken[s] = 1;
    for (size_t h = 0; h < queue.size() && (int64_t)queue.size() < limit; ++h) {
      const int64_t v = queue[h];
      for (int64_t j = g.ptr[v]; j < g.ptr[v + 1] && (int64_t)queue.size() < limit; ++j) {
        const int64_t w = g.adj[j];
        if (!taken[w]) {
          taken[w] = 1;
          queue.push_back(w);
        }
      }
    }
    for (int64_t v : queue) order[out++] = v;
  }
  ESR_HREQUIRE(out == n, "esr_block_order: internal count %lld != %lld", (long long)out,
               (long long)n);
  return ESR_OK;
  ESR_HGUARD_END
}

ESR_HAPI int esr_bandwidth(const int64_t *a, const int64_t *b, int64_t m, int64_t *bw) {
  ESR_HREQUIRE(bw != nullptr && (m == 0 || (a && b)), "esr_bandwidth: null pointer");
  int64_t w = 0;
  for (int64_t k = 0; k < m; ++k) w = std::max(w, a[k] > b[k] ? a[k] - b[k] : b[k] - a[k]);
  *bw = w;
  return ESR_OK;
}
