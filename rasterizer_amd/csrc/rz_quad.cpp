// QuadDecomposition::decompose (QuadDecomposition.cpp:311-445) on the host:
// pairs a triangle mesh's triangles into quads along shared edges whose two
// triangles lie within 0.5 units of each other's plane, as a maximum matching
// of the candidate graph (greedy start, then Edmonds' augmenting paths with
// blossom shrinking).  The output must equal the reference's quad for quad,
// so the candidate graph's adjacency order, the greedy pass and every step of
// the augmenting-path search follow the reference's order; the data
// structures are this file's own (sorted half-edge groups instead of a hash
// map of vectors, flat arrays, an explicit stack for path compression).
#include <immintrin.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <deque>
#include <new>
#include <utility>
#include <vector>

#include "rasterizer_amd.h"
#include "rz_quad.h"

namespace {

// --- candidate test (QuadDecomposition.cpp:326-343) ------------------------
// VectorMath.h's cross product in its lane order: (a * b.yzx - a.yzx * b).yzx.
inline __m128 cross3(__m128 a, __m128 b) {
  const __m128 ayzx = _mm_shuffle_ps(a, a, _MM_SHUFFLE(3, 0, 2, 1));
  const __m128 byzx = _mm_shuffle_ps(b, b, _MM_SHUFFLE(3, 0, 2, 1));
  const __m128 c = _mm_sub_ps(_mm_mul_ps(a, byzx), _mm_mul_ps(ayzx, b));
  return _mm_shuffle_ps(c, c, _MM_SHUFFLE(3, 0, 2, 1));
}
// x * RSQRTPS(dot3(x, x)) (VectorMath.h:20-23), with this CPU's RSQRTPS.
inline __m128 unit(__m128 v) { return _mm_mul_ps(v, _mm_rsqrt_ps(_mm_dp_ps(v, v, 0x7F))); }

// Both triangles (v0, v1, v2) and (v2, v3, v0) keep v1 - v3 within 0.5 of
// their planes; a NaN distance does not exceed the bound (COMISS is false).
bool mergeable(__m128 v0, __m128 v1, __m128 v2, __m128 v3) {
  const __m128 sign = _mm_set1_ps(-0.0f);
  const __m128 n0 = unit(cross3(_mm_sub_ps(v1, v0), _mm_sub_ps(v2, v0)));
  const __m128 n2 = unit(cross3(_mm_sub_ps(v3, v2), _mm_sub_ps(v0, v2)));
  const __m128 d = _mm_sub_ps(v1, v3);
  const __m128 a = _mm_andnot_ps(sign, _mm_dp_ps(n0, d, 0x7F));
  const __m128 b = _mm_andnot_ps(sign, _mm_dp_ps(n2, d, 0x7F));
  const __m128 lim = _mm_set1_ps(0.5f);
  return !(_mm_comigt_ss(a, lim) || _mm_comigt_ss(b, lim));
}

// --- directed edges ---------------------------------------------------------
// Half-edge k of triangle t is (i[k], i[k+1 mod 3]) with apex i[k+2 mod 3];
// its insertion number is 3t + k, the order in which the reference appends it
// to its edge's list (QuadDecomposition.cpp:361-363).  Sorting by (edge,
// insertion number) gives each directed edge's list in that order.
struct HalfEdge {
  uint64_t key;  // from << 32 | to
  uint32_t seq;  // 3t + k
  uint32_t apex;
};

struct EdgeIndex {
  std::vector<HalfEdge> e;
  // Index of the first half-edge with this key, or e.size().
  size_t first(uint64_t key) const {
    auto it = std::lower_bound(e.begin(), e.end(), key, [](const HalfEdge& h, uint64_t k) { return h.key < k; });
    return size_t(it - e.begin());
  }
};

inline uint64_t edge_key(uint32_t a, uint32_t b) { return (uint64_t(a) << 32) | b; }

// --- maximum matching (QuadDecomposition.cpp:28-309) ------------------------
class Matcher {
 public:
  Matcher(int n, const uint32_t* adjStart, const int* adj)
      : start_(adjStart), adj_(adj), n_(n), mate_(size_t(n_), -1), bridge_(size_t(n_)),
        depth_(size_t(n_), 0), parent_(size_t(n_), -1), blossom_(size_t(n_), 0), token_(size_t(n_), 0) {}

  // Greedy maximal matching in vertex and adjacency order (:34-54), then an
  // augmenting-path search from each vertex the greedy pass left single
  // (:56-67).
  void run() {
    std::vector<int> single;
    for (int v = 0; v < n_; ++v) {
      if (mate_[v] != -1) continue;
      bool found = false;
      for (uint32_t k = start_[v]; k < start_[v + 1]; ++k) {
        const int w = adj_[k];
        if (mate_[w] == -1) {
          pair(v, w);
          found = true;
          break;
        }
      }
      if (!found) single.push_back(v);
    }
    std::vector<int> path;
    for (int v : single) {
      if (mate_[v] != -1) continue;
      if (search(v, path)) {
        for (size_t i = 0; i + 1 < path.size(); i += 2) pair(path[i], path[i + 1]);
        path.clear();
      }
    }
  }

  int mate(int v) const { return mate_[v]; }

 private:
  void pair(int v, int w) {
    mate_[v] = w;
    mate_[w] = v;
  }

  // Representative of x's blossom in the current search; a vertex outside
  // the search tree is its own (:277-290).  Compresses the path.
  int find(int x) {
    int r = x;
    while (token_[r] == round_ && blossom_[r] != r) r = blossom_[r];
    while (token_[x] == round_ && blossom_[x] != x) {
      const int next = blossom_[x];
      blossom_[x] = r;
      x = next;
    }
    return r;
  }
  void unite(int x, int y) {  // :265-269
    const int rx = find(x);
    blossom_[rx] = find(y);
  }
  void represent(int x) {  // :271-276
    const int rx = find(x);
    blossom_[rx] = x;
    blossom_[x] = x;
  }

  // Breadth-first growth of the alternating tree from root (:94-126).
  bool search(int root, std::vector<int>& path) {
    ++round_;
    depth_[root] = 0;
    parent_[root] = -1;
    token_[root] = round_;
    blossom_[root] = root;
    queue_.clear();
    queue_.push_back(root);
    while (!queue_.empty()) {
      const int v = queue_.front();
      queue_.pop_front();
      for (uint32_t k = start_[v]; k < start_[v + 1]; ++k) {
        if (examine(root, v, adj_[k], path)) {
          queue_.clear();
          return true;
        }
      }
    }
    return false;
  }

  // One edge of the tree's frontier (:128-155).
  bool examine(int root, int v, int w, std::vector<int>& path) {
    const int vb = find(v);
    const int wb = find(w);
    if (vb == wb) return false;
    if (token_[wb] != round_) {
      if (mate_[w] == -1) {
        path.push_back(w);  // :157-161
        trace(v, root, path);
        return true;
      }
      grow(v, w);
    } else if (depth_[wb] % 2 == 0) {
      shrink(v, w);
    }
    return false;
  }

  // w (unmatched in the tree) and its mate u join below v (:163-182).
  void grow(int v, int w) {
    const int u = mate_[w];
    depth_[w] = depth_[v] + 1 + (depth_[v] & 1);
    parent_[w] = v;
    token_[w] = round_;
    blossom_[w] = w;
    depth_[u] = depth_[w] + 1;
    parent_[u] = w;
    token_[u] = round_;
    blossom_[u] = u;
    queue_.push_back(u);
  }

  // The odd cycle closed by edge (v, w) becomes one blossom (:184-221).
  void shrink(int v, int w) {
    int a = v, b = w;
    while (b != a) {
      if (depth_[a] > depth_[b]) a = parent_[a];
      else b = parent_[b];
    }
    const int base = find(a);
    shrink_side(base, v, w);
    shrink_side(base, w, v);
  }
  void shrink_side(int base, int v, int w) {
    int u = find(v);
    while (u != base) {
      unite(base, u);
      u = mate_[u];
      unite(base, u);
      represent(base);
      queue_.push_back(u);
      bridge_[u] = {v, w};
      u = find(parent_[u]);
    }
  }

  // Appends the tree path from s to t, expanding blossoms through their
  // bridges (:223-252).
  void trace(int s, int t, std::vector<int>& path) {
    while (s != t) {
      if (depth_[s] % 2 == 0) {
        path.push_back(s);
        path.push_back(mate_[s]);
        s = parent_[mate_[s]];
        continue;
      }
      const std::pair<int, int> br = bridge_[s];
      path.push_back(s);
      const size_t from = path.size();
      trace(br.first, mate_[s], path);
      std::reverse(path.begin() + std::ptrdiff_t(from), path.end());
      s = br.second;
    }
    path.push_back(s);
  }

  const uint32_t* start_;
  const int* adj_;
  int n_;
  std::vector<int> mate_;
  std::vector<std::pair<int, int>> bridge_;
  std::vector<int> depth_, parent_, blossom_, token_;
  int round_ = 0;
  std::deque<int> queue_;
};

}  // namespace

namespace rzq {

void match(uint32_t n, const uint32_t* start, const int* adj, int* mate) {
  Matcher m(int(n), start, adj);
  m.run();
  for (uint32_t v = 0; v < n; ++v) mate[v] = m.mate(int(v));
}

}  // namespace rzq

extern "C" int rz_quad_decompose(const uint32_t* indices, size_t nIndices, const float* vertices, size_t nVertices,
                                 uint32_t* out, size_t* nOut) {
  if (!nOut || (nIndices >= 3 && (!indices || !vertices || !out))) return RZ_ERR_INVALID_ARGUMENT;
  *nOut = 0;
  const size_t nTri = nIndices / 3;  // a trailing partial triangle is ignored, as the reference's size / 3
  if (nTri == 0) return RZ_OK;
  if (nTri > size_t(0x7FFFFFFF) / 3 || nVertices > 0xFFFFFFFFull) return RZ_ERR_TOO_LARGE;
  for (size_t i = 0; i < 3 * nTri; ++i)
    if (indices[i] >= nVertices) return RZ_ERR_INVALID_ARGUMENT;
  try {
    const uint32_t T = uint32_t(nTri);
    EdgeIndex ex;
    ex.e.resize(size_t(3) * T);
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t* i = indices + 3 * size_t(t);
      for (uint32_t k = 0; k < 3; ++k)
        ex.e[3 * size_t(t) + k] = HalfEdge{edge_key(i[k], i[(k + 1) % 3]), 3 * t + k, i[(k + 2) % 3]};
    }
    std::sort(ex.e.begin(), ex.e.end(),
              [](const HalfEdge& a, const HalfEdge& b) { return a.key != b.key ? a.key < b.key : a.seq < b.seq; });
    // For each half-edge (by insertion number), the first entry of its
    // reverse edge's list.
    std::vector<uint32_t> reverseFirst(size_t(3) * T);
    for (const HalfEdge& h : ex.e) {
      const uint64_t rev = (h.key << 32) | (h.key >> 32);
      reverseFirst[h.seq] = uint32_t(ex.first(rev));
    }
    auto vert = [&](uint32_t v) { return _mm_loadu_ps(vertices + 4 * size_t(v)); };

    // Candidate graph (:355-384): triangle t meets, across each of its edges
    // in order, the triangles already listed on the reverse edge (its own
    // included), in list order; a mergeable pair is linked both ways.  Links
    // are appended in that global order, so a count pass and a fill pass
    // over the same loop give flat adjacency lists in the reference's order.
    std::vector<uint32_t> degree(size_t(T) + 1, 0u);
    std::vector<uint32_t> linkA, linkB;  // link j joins linkA[j] -> linkB[j] (and back)
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t* i = indices + 3 * size_t(t);
      for (uint32_t k = 0; k < 3; ++k) {
        const uint64_t rev = edge_key(i[(k + 1) % 3], i[k]);
        for (size_t j = reverseFirst[3 * t + k]; j < ex.e.size() && ex.e[j].key == rev && ex.e[j].seq <= 3 * t + 2;
             ++j) {
          const uint32_t nb = ex.e[j].seq / 3;
          if (mergeable(vert(i[k]), vert(ex.e[j].apex), vert(i[(k + 1) % 3]), vert(i[(k + 2) % 3]))) {
            linkA.push_back(t);
            linkB.push_back(nb);
            ++degree[t];
            ++degree[nb];
          }
        }
      }
    }
    std::vector<uint32_t> start(size_t(T) + 1, 0u);
    for (uint32_t t = 0; t < T; ++t) start[t + 1] = start[t] + degree[t];
    std::vector<int> adj(start[T]);
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (size_t j = 0; j < linkA.size(); ++j) {
      adj[fill[linkA[j]]++] = int(linkB[j]);
      adj[fill[linkB[j]]++] = int(linkA[j]);
    }
    std::vector<uint32_t>().swap(linkA);
    std::vector<uint32_t>().swap(linkB);

    Matcher m(int(T), start.data(), adj.data());
    m.run();

    // Quads in triangle order (:392-441): a single triangle becomes the
    // degenerate quad (i0, i2, i1, i0); a matched pair is written once, by
    // its lower triangle, from the first of its edges whose reverse list
    // holds the partner.  A triangle matched to itself writes nothing.
    size_t w = 0;
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t* i = indices + 3 * size_t(t);
      const int nb = m.mate(int(t));
      if (nb == -1) {
        out[w++] = i[0];
        out[w++] = i[2];
        out[w++] = i[1];
        out[w++] = i[0];
        continue;
      }
      if (t >= uint32_t(nb)) continue;
      bool done = false;
      for (uint32_t k = 0; k < 3 && !done; ++k) {
        const uint64_t rev = edge_key(i[(k + 1) % 3], i[k]);
        for (size_t j = reverseFirst[3 * t + k]; j < ex.e.size() && ex.e[j].key == rev; ++j) {
          if (ex.e[j].seq / 3 == uint32_t(nb)) {
            out[w++] = i[k];
            out[w++] = i[(k + 2) % 3];
            out[w++] = i[(k + 1) % 3];
            out[w++] = ex.e[j].apex;
            done = true;
            break;
          }
        }
      }
    }
    *nOut = w;
  } catch (const std::bad_alloc&) {
    return RZ_ERR_OUT_OF_MEMORY;
  }
  return RZ_OK;
}
