// Host CSR graph: reference-format I/O, validation, generators, DAG.
#include "host_graph.hpp"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <sstream>

#include "mgpm.h"

namespace mgpm {

int32_t CsrGraph::max_degree() const {
  int32_t md = 0;
#pragma omp parallel for reduction(max : md) schedule(static)
  for (int32_t v = 0; v < n; ++v) md = std::max(md, degree(v));
  return md;
}

void prefix_sum(const std::vector<int64_t> &in, std::vector<int64_t> &out) {
  const size_t len = in.size();
  out.resize(len + 1);
  const int nt = omp_get_max_threads();
  std::vector<int64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num();
    const int T = omp_get_num_threads();
    const size_t lo = len * t / T, hi = len * (t + 1) / T;
    int64_t s = 0;
    for (size_t i = lo; i < hi; ++i) s += in[i];
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int k = 1; k <= T; ++k) part[k] += part[k - 1];
    int64_t run = part[t];
    for (size_t i = lo; i < hi; ++i) {
      out[i] = run;
      run += in[i];
    }
  }
  int64_t total = 0;
  // part[T] holds the grand total; recompute cheaply for the tail entry.
  if (len) total = out[len - 1] + in[len - 1];
  out[len] = total;
}

// ---------------------------------------------------------------- file I/O

namespace {

bool file_size(const std::string &path, uint64_t &sz) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long long s = std::ftell(f);
  std::fclose(f);
  if (s < 0) return false;
  sz = static_cast<uint64_t>(s);
  return true;
}

int read_exact(const std::string &path, void *dst, uint64_t bytes,
               std::string &err) {
  uint64_t sz = 0;
  if (!file_size(path, sz)) {
    err = "Failed to open file: " + path;
    return MGPM_ERR_IO;
  }
  if (sz < bytes) {
    err = "file too short: " + path + " has " + std::to_string(sz) +
          " bytes, expected " + std::to_string(bytes);
    return MGPM_ERR_IO;
  }
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "Failed to open file: " + path;
    return MGPM_ERR_IO;
  }
  uint64_t got = bytes ? std::fread(dst, 1, bytes, f) : 0;
  std::fclose(f);
  if (got != bytes) {
    err = "short read: " + path;
    return MGPM_ERR_IO;
  }
  return MGPM_OK;
}

int write_exact(const std::string &path, const void *src, uint64_t bytes,
                std::string &err) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot create file: " + path;
    return MGPM_ERR_IO;
  }
  uint64_t put = bytes ? std::fwrite(src, 1, bytes, f) : 0;
  int rc = std::fclose(f);
  if (put != bytes || rc != 0) {
    err = "short write: " + path;
    return MGPM_ERR_IO;
  }
  return MGPM_OK;
}

}  // namespace

namespace {

bool file_exists(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

// Label i of `classes` for a missing label file: 1..classes.
uint32_t fallback_label(uint64_t i, uint32_t classes, uint64_t salt) {
  uint64_t z = i + salt + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<uint32_t>(z % classes) + 1u;
}

}  // namespace

int check_labels(const CsrGraph &g, std::string &err, bool check_vertex,
                 bool check_edge) {
  if (check_vertex && !g.vlabels.empty()) {
    if (!(g.num_vertex_classes > 0 && g.num_vertex_classes < 255)) {
      err = "vertex labels need 0 < num_vertex_classes < 255";
      return MGPM_ERR_FORMAT;
    }
    if (static_cast<int64_t>(g.vlabels.size()) != g.n) {
      err = "vertex labels: one per vertex";
      return MGPM_ERR_FORMAT;
    }
    bool seen[256] = {};
    int distinct = 0;
    for (uint8_t l : g.vlabels)
      if (!seen[l]) {
        seen[l] = true;
        ++distinct;
      }
    if (distinct != g.num_vertex_classes) {
      err = "vertex labels: " + std::to_string(distinct) +
            " distinct labels, num_vertex_classes says " +
            std::to_string(g.num_vertex_classes);
      return MGPM_ERR_FORMAT;
    }
  }
  if (check_edge && !g.elabels.empty()) {
    if (!(g.num_edge_classes > 0 && g.num_edge_classes < 65535)) {
      err = "edge labels need 0 < num_edge_classes < 65535";
      return MGPM_ERR_FORMAT;
    }
    if (static_cast<int64_t>(g.elabels.size()) != g.nnz) {
      err = "edge labels: one per CSR entry";
      return MGPM_ERR_FORMAT;
    }
    std::vector<uint8_t> seen(65536, 0);
    int distinct = 0;
    for (uint16_t l : g.elabels)
      if (!seen[l]) {
        seen[l] = 1;
        ++distinct;
      }
    if (distinct > g.num_edge_classes) {
      err = "edge labels: " + std::to_string(distinct) +
            " distinct labels, more than num_edge_classes " +
            std::to_string(g.num_edge_classes);
      return MGPM_ERR_FORMAT;
    }
  }
  return MGPM_OK;
}

std::vector<std::pair<int32_t, int32_t>> nlf(const CsrGraph &g, int32_t v) {
  int32_t cnt[256] = {};
  for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e)
    ++cnt[g.vlabels[g.colidx[e]]];
  std::vector<std::pair<int32_t, int32_t>> out;
  for (int32_t l = 0; l < 256; ++l)
    if (cnt[l]) out.emplace_back(l, cnt[l]);
  return out;
}

void nlf_csr(const CsrGraph &g, int64_t *offsets, int32_t *labels,
             int32_t *counts) {
  const int32_t n = g.n;
  if (!labels || !counts) {
    offsets[0] = 0;
#pragma omp parallel
    {
      uint32_t stamp[256] = {};
#pragma omp for schedule(dynamic, 4096)
      for (int32_t v = 0; v < n; ++v) {
        const uint32_t tag = static_cast<uint32_t>(v) + 1;
        int64_t k = 0;
        for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e) {
          const uint8_t l = g.vlabels[g.colidx[e]];
          if (stamp[l] != tag) {
            stamp[l] = tag;
            ++k;
          }
        }
        offsets[v + 1] = k;
      }
    }
    for (int32_t v = 0; v < n; ++v) offsets[v + 1] += offsets[v];
    return;
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < n; ++v) {
    int32_t cnt[256] = {};
    for (int64_t e = g.rowptr[v]; e < g.rowptr[v + 1]; ++e)
      ++cnt[g.vlabels[g.colidx[e]]];
    int64_t o = offsets[v];
    for (int32_t l = 0; l < 256; ++l)
      if (cnt[l]) {
        labels[o] = l;
        counts[o++] = cnt[l];
      }
  }
}

bool label_frequency(const CsrGraph &g, std::vector<int64_t> &freq) {
  freq.assign(static_cast<size_t>(g.num_vertex_classes) + 1, 0);
  for (uint8_t l : g.vlabels) {
    if (l >= freq.size()) return false;
    ++freq[l];
  }
  return true;
}

int read_graph(const std::string &prefix, CsrGraph &g, std::string &err,
               int flags) {
  std::ifstream meta(prefix + ".meta.txt");
  if (!meta) {
    err = "Failed to open file: " + prefix + ".meta.txt";
    return MGPM_ERR_IO;
  }
  long long f[10];
  for (int i = 0; i < 10; ++i) {
    if (!(meta >> f[i])) {
      err = "graph.meta.txt: expected 10 integers, got " + std::to_string(i);
      return MGPM_ERR_FORMAT;
    }
  }
  const long long n = f[0], nnz = f[1], md = f[6];
  // graph.cc:24-28 asserts these sizes and 0 < max_degree < n.
  if (f[2] != 4 || f[3] != 8 || f[4] != 1 || f[5] != 2) {
    err = "graph.meta.txt: type sizes must be vid=4 eid=8 vlabel=1 elabel=2";
    return MGPM_ERR_FORMAT;
  }
  if (n <= 0 || n > INT32_MAX || nnz < 0) {
    err = "graph.meta.txt: bad vertex/edge count";
    return MGPM_ERR_FORMAT;
  }
  if (!(md > 0 && md < n)) {
    err = "graph.meta.txt: max_degree must satisfy 0 < max_degree < n";
    return MGPM_ERR_FORMAT;
  }
  g.n = static_cast<int32_t>(n);
  g.nnz = nnz;
  g.meta_max_degree = static_cast<int32_t>(md);
  const long long nvc = f[8], nec = f[9];
  try {
    g.rowptr.assign(static_cast<size_t>(n) + 1, 0);
    g.colidx.assign(static_cast<size_t>(nnz), 0);
  } catch (const std::bad_alloc &) {
    err = "out of host memory reading graph";
    return MGPM_ERR_NOMEM;
  }
  int rc = read_exact(prefix + ".vertex.bin", g.rowptr.data(),
                      (static_cast<uint64_t>(n) + 1) * 8, err);
  if (rc) return rc;
  rc = read_exact(prefix + ".edge.bin", g.colidx.data(),
                  static_cast<uint64_t>(nnz) * 4, err);
  if (rc) return rc;
  if ((rc = check_simple(g, err))) return rc;
  bool vfile = false, efile = false;
  if (flags & kLoadVertexLabels) {
    if (!(nvc > 0 && nvc < 255)) {  // graph.cc:40-41
      err = "graph.meta.txt: vertex labels need 0 < num_vertex_classes < 255";
      return MGPM_ERR_FORMAT;
    }
    g.num_vertex_classes = static_cast<int32_t>(nvc);
    g.vlabels.assign(static_cast<size_t>(n), 0);
    const std::string path = prefix + ".vlabel.bin";
    if (file_exists(path)) {
      if ((rc = read_exact(path, g.vlabels.data(), static_cast<uint64_t>(n),
                           err)))
        return rc;
      vfile = true;
    } else {
      for (int64_t v = 0; v < n; ++v)
        g.vlabels[v] = static_cast<uint8_t>(
            fallback_label(static_cast<uint64_t>(v), static_cast<uint32_t>(nvc), 1));
    }
  }
  if (flags & kLoadEdgeLabels) {
    if (!(nec > 0 && nec < 65535)) {  // graph.cc:62-63
      err = "graph.meta.txt: edge labels need 0 < num_edge_classes < 65535";
      return MGPM_ERR_FORMAT;
    }
    g.num_edge_classes = static_cast<int32_t>(nec);
    g.elabels.assign(static_cast<size_t>(nnz), 0);
    const std::string path = prefix + ".elabel.bin";
    if (file_exists(path)) {
      if ((rc = read_exact(path, g.elabels.data(),
                           static_cast<uint64_t>(nnz) * 2, err)))
        return rc;
      efile = true;
    } else {
      for (int64_t e = 0; e < nnz; ++e)
        g.elabels[e] = static_cast<uint16_t>(
            fallback_label(static_cast<uint64_t>(e), static_cast<uint32_t>(nec), 2));
    }
  }
  // generated labels are not checked against the class counts (graph.cc
  // checks only labels read from a file)
  return check_labels(g, err, vfile, efile);
}

int write_graph(const std::string &prefix, const CsrGraph &g,
                std::string &err) {
  {
    std::ofstream meta(prefix + ".meta.txt");
    if (!meta) {
      err = "cannot create file: " + prefix + ".meta.txt";
      return MGPM_ERR_IO;
    }
    meta << g.n << " " << g.nnz << " 4 8 1 2 " << g.max_degree() << " 0 "
         << (g.vlabels.empty() ? 0 : g.num_vertex_classes) << " "
         << (g.elabels.empty() ? 0 : g.num_edge_classes) << "\n";
    if (!meta) {
      err = "short write: " + prefix + ".meta.txt";
      return MGPM_ERR_IO;
    }
  }
  int rc = write_exact(prefix + ".vertex.bin", g.rowptr.data(),
                       (static_cast<uint64_t>(g.n) + 1) * 8, err);
  if (rc) return rc;
  if ((rc = write_exact(prefix + ".edge.bin", g.colidx.data(),
                        static_cast<uint64_t>(g.nnz) * 4, err)))
    return rc;
  if (!g.vlabels.empty() &&
      (rc = write_exact(prefix + ".vlabel.bin", g.vlabels.data(),
                        static_cast<uint64_t>(g.n), err)))
    return rc;
  if (!g.elabels.empty() &&
      (rc = write_exact(prefix + ".elabel.bin", g.elabels.data(),
                        static_cast<uint64_t>(g.nnz) * 2, err)))
    return rc;
  return MGPM_OK;
}

int check_simple(const CsrGraph &g, std::string &err) {
  if (g.n < 0 || static_cast<int64_t>(g.rowptr.size()) != int64_t(g.n) + 1 ||
      static_cast<int64_t>(g.colidx.size()) != g.nnz) {
    err = "CSR arrays have inconsistent sizes";
    return MGPM_ERR_GRAPH;
  }
  if (g.rowptr[0] != 0 || g.rowptr[g.n] != g.nnz) {
    err = "row offsets must start at 0 and end at n_edges";
    return MGPM_ERR_GRAPH;
  }
  std::atomic<int> bad{0};
  std::atomic<int64_t> where{-1};
#pragma omp parallel for schedule(dynamic, 4096)
  for (int32_t v = 0; v < g.n; ++v) {
    if (bad.load(std::memory_order_relaxed)) continue;
    const int64_t b = g.rowptr[v], e = g.rowptr[v + 1];
    if (b > e || b < 0 || e > g.nnz) {
      bad = 1;
      where = v;
      continue;
    }
    int32_t prev = -1;
    for (int64_t i = b; i < e; ++i) {
      const int32_t u = g.colidx[i];
      int code = 0;
      if (u < 0 || u >= g.n) code = 2;
      else if (u == v) code = 3;
      else if (u <= prev) code = 4;
      if (code) {
        bad = code;
        where = v;
        break;
      }
      prev = u;
    }
  }
  if (!bad) {
    // symmetry: u in N(v) => v in N(u)
#pragma omp parallel for schedule(dynamic, 4096)
    for (int32_t v = 0; v < g.n; ++v) {
      if (bad.load(std::memory_order_relaxed)) continue;
      for (int64_t i = g.rowptr[v]; i < g.rowptr[v + 1]; ++i) {
        const int32_t u = g.colidx[i];
        const int32_t *lo = g.colidx.data() + g.rowptr[u];
        const int32_t *hi = g.colidx.data() + g.rowptr[u + 1];
        if (!std::binary_search(lo, hi, v)) {
          bad = 5;
          where = v;
          break;
        }
      }
    }
  }
  static const char *what[] = {"",
                               "row offsets not monotone or out of range",
                               "neighbour id out of range",
                               "self-loop",
                               "row not strictly increasing (unsorted or duplicate)",
                               "adjacency not symmetric"};
  if (bad) {
    err = std::string("invalid CSR: ") + what[bad.load()] + " at vertex " +
          std::to_string(where.load());
    return MGPM_ERR_GRAPH;
  }
  return MGPM_OK;
}

// ---------------------------------------------------------------- builders

void csr_from_draws(int32_t n, std::vector<uint64_t> &pairs, CsrGraph &g) {
  const int64_t nd = static_cast<int64_t>(pairs.size());
  std::vector<std::atomic<int64_t>> cnt(static_cast<size_t>(n));
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < n; ++v) cnt[v].store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nd; ++i) {
    const int32_t s = static_cast<int32_t>(pairs[i] >> 32);
    const int32_t d = static_cast<int32_t>(pairs[i] & 0xffffffffu);
    if (s == d) continue;
    cnt[s].fetch_add(1, std::memory_order_relaxed);
    cnt[d].fetch_add(1, std::memory_order_relaxed);
  }
  std::vector<int64_t> deg(n), off;
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < n; ++v) deg[v] = cnt[v].load();
  prefix_sum(deg, off);
  std::vector<int32_t> tmp(static_cast<size_t>(off[n]));
#pragma omp parallel for schedule(static)
  for (int32_t v = 0; v < n; ++v) cnt[v].store(off[v]);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nd; ++i) {
    const int32_t s = static_cast<int32_t>(pairs[i] >> 32);
    const int32_t d = static_cast<int32_t>(pairs[i] & 0xffffffffu);
    if (s == d) continue;
    tmp[cnt[s].fetch_add(1, std::memory_order_relaxed)] = d;
    tmp[cnt[d].fetch_add(1, std::memory_order_relaxed)] = s;
  }
  std::vector<uint64_t>().swap(pairs);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; ++v) {
    int32_t *b = tmp.data() + off[v], *e = tmp.data() + off[v + 1];
    std::sort(b, e);
    deg[v] = std::unique(b, e) - b;
  }
  g.n = n;
  prefix_sum(deg, g.rowptr);
  g.nnz = g.rowptr[n];
  g.colidx.assign(static_cast<size_t>(g.nnz), 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t v = 0; v < n; ++v)
    std::copy(tmp.data() + off[v], tmp.data() + off[v] + deg[v],
              g.colidx.data() + g.rowptr[v]);
  g.meta_max_degree = 0;
}

namespace {
inline uint64_t mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xbf58476d1ce4e5b9ULL;
  z ^= z >> 27;
  z *= 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return z;
}
// j-th 64-bit random word of draw i under `key` (counter-based).
inline uint64_t draw_word(uint64_t key, uint64_t i, uint64_t j) {
  return mix64(key + (i * 32 + j + 1) * 0x9E3779B97F4A7C15ULL);
}
}  // namespace

void generate_rmat(int scale, int edge_factor, double a, double b, double c,
                   uint64_t seed, CsrGraph &g) {
  const int32_t n = int32_t(1) << scale;
  const int64_t draws = int64_t(edge_factor) << scale;
  const double two32 = 4294967296.0;
  const uint64_t tA = static_cast<uint64_t>(a * two32);
  const uint64_t tAB = static_cast<uint64_t>((a + b) * two32);
  const uint64_t tABC = static_cast<uint64_t>((a + b + c) * two32);
  const uint64_t key = mix64(seed ^ 0x5851F42D4C957F2DULL);
  std::vector<uint64_t> pairs(static_cast<size_t>(draws));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < draws; ++i) {
    uint32_t s = 0, d = 0;
    uint64_t w = 0;
    for (int l = 0; l < scale; ++l) {
      if ((l & 1) == 0) w = draw_word(key, static_cast<uint64_t>(i), l >> 1);
      const uint64_t r = (l & 1) ? (w >> 32) : (w & 0xffffffffULL);
      const uint32_t bit = 1u << (scale - 1 - l);
      if (r < tA) {
      } else if (r < tAB) {
        d |= bit;
      } else if (r < tABC) {
        s |= bit;
      } else {
        s |= bit;
        d |= bit;
      }
    }
    pairs[i] = (uint64_t(s) << 32) | d;
  }
  csr_from_draws(n, pairs, g);
}

void generate_er(int32_t n, int64_t draws, uint64_t seed, CsrGraph &g) {
  const uint64_t key = mix64(seed ^ 0x2545F4914F6CDD1DULL);
  std::vector<uint64_t> pairs(static_cast<size_t>(draws));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < draws; ++i) {
    const uint64_t w = draw_word(key, static_cast<uint64_t>(i), 0);
    const uint64_t s = ((w >> 32) * static_cast<uint64_t>(n)) >> 32;
    const uint64_t d = ((w & 0xffffffffULL) * static_cast<uint64_t>(n)) >> 32;
    pairs[i] = (s << 32) | d;
  }
  csr_from_draws(n, pairs, g);
}

// ---------------------------------------------------------------- DAG

void build_dag(const CsrGraph &g, Dag &d) {
  const int32_t n = g.n;
  std::vector<int64_t> od(n);
  d.out_deg.assign(n, 0);
  int32_t mx = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : mx)
  for (int32_t s = 0; s < n; ++s) {
    const int32_t ds = g.degree(s);
    int32_t c = 0;
    for (int64_t i = g.rowptr[s]; i < g.rowptr[s + 1]; ++i) {
      const int32_t t = g.colidx[i];
      const int32_t dt = g.degree(t);
      c += (dt > ds || (dt == ds && t > s));
    }
    d.out_deg[s] = c;
    od[s] = c;
    mx = std::max(mx, c);
  }
  d.max_out = mx;
  prefix_sum(od, d.rowptr);
  d.colidx.assign(static_cast<size_t>(d.rowptr[n]), 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t s = 0; s < n; ++s) {
    const int32_t ds = g.degree(s);
    int64_t o = d.rowptr[s];
    for (int64_t i = g.rowptr[s]; i < g.rowptr[s + 1]; ++i) {
      const int32_t t = g.colidx[i];
      const int32_t dt = g.degree(t);
      if (dt > ds || (dt == ds && t > s)) d.colidx[o++] = t;
    }
  }
}

void build_division(const CsrGraph &g, const Dag &d, int32_t deg_thd,
                    Division &out) {
  const int32_t n = g.n;
  std::vector<int64_t> sd(n, 0), ld(n, 0);
  const auto &nd = d.out_deg;
  // graph.cc:237-312: both loops walk the undirected adjacency.
  auto classify = [&](int32_t src, int32_t dst) -> int {
    const int32_t a = nd[src], b = nd[dst];
    if (std::min(a, b) < 1) return 0;
    if (std::max(a, b) <= deg_thd) {
      const int32_t ds = g.degree(src), dt = g.degree(dst);
      return (dt > ds || (dt == ds && dst > src)) ? 1 : 0;
    }
    return (a > b || (a == b && src > dst)) ? 2 : 0;
  };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t s = 0; s < n; ++s) {
    for (int64_t i = g.rowptr[s]; i < g.rowptr[s + 1]; ++i) {
      const int k = classify(s, g.colidx[i]);
      sd[s] += (k == 1);
      ld[s] += (k == 2);
    }
  }
  prefix_sum(sd, out.small_rowptr);
  prefix_sum(ld, out.large_rowptr);
  out.small_colidx.assign(static_cast<size_t>(out.small_rowptr[n]), 0);
  out.large_colidx.assign(static_cast<size_t>(out.large_rowptr[n]), 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int32_t s = 0; s < n; ++s) {
    int64_t so = out.small_rowptr[s], lo = out.large_rowptr[s];
    for (int64_t i = g.rowptr[s]; i < g.rowptr[s + 1]; ++i) {
      const int32_t t = g.colidx[i];
      const int k = classify(s, t);
      if (k == 1) out.small_colidx[so++] = t;
      else if (k == 2) out.large_colidx[lo++] = t;
    }
  }
}

namespace {

// Next unsigned decimal integer in [p, e) on the current line; skips spaces
// and tabs only.  False if the line has no further number.
bool next_int(const char *&p, const char *e, int64_t &v) {
  while (p < e && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  if (p >= e || *p < '0' || *p > '9') return false;
  int64_t x = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    x = x * 10 + (*p - '0');
    if (x > (int64_t(1) << 40)) return false;
    ++p;
  }
  v = x;
  return true;
}

}  // namespace

int read_edge_file(const std::string &path, int format, int base, CsrGraph &g,
                   std::string &err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) {
    err = "Failed to open file: " + path;
    return MGPM_ERR_IO;
  }
  std::string text((std::istreambuf_iterator<char>(in)),
                   std::istreambuf_iterator<char>());
  const char *p = text.data(), *end = p + text.size();
  int64_t line_no = 0;
  int64_t n_decl = -1, entries = -1;
  int64_t max_id = -1;
  std::vector<uint64_t> pairs;
  bool banner = false;
  while (p < end) {
    const char *eol = static_cast<const char *>(memchr(p, '\n', end - p));
    if (!eol) eol = end;
    ++line_no;
    const char *q = p;
    p = eol + 1;
    while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q >= eol) continue;  // blank line
    if (format == 1 && line_no == 1) {
      banner = (eol - q) >= 14 && std::string(q, 14) == "%%MatrixMarket";
      if (!banner) {
        err = path + ": missing %%MatrixMarket banner";
        return MGPM_ERR_FORMAT;
      }
      continue;
    }
    if (*q == '%' || (format == 2 && *q == '#')) continue;
    int64_t a = 0, b = 0, c = 0;
    if (!next_int(q, eol, a) || !next_int(q, eol, b)) {
      err = path + ":" + std::to_string(line_no) + ": expected two vertex ids";
      return MGPM_ERR_FORMAT;
    }
    if (format == 1 && n_decl < 0) {  // the size line
      if (!next_int(q, eol, c)) {
        err = path + ":" + std::to_string(line_no) + ": expected rows cols entries";
        return MGPM_ERR_FORMAT;
      }
      n_decl = std::max(a, b);
      entries = c;
      pairs.reserve(static_cast<size_t>(std::min<int64_t>(c, int64_t(1) << 32)));
      continue;
    }
    a -= base;
    b -= base;
    const int64_t lim = format == 1 ? n_decl : (int64_t(1) << 31) - 1;
    if (a < 0 || b < 0 || a >= lim || b >= lim) {
      err = path + ":" + std::to_string(line_no) + ": vertex id out of range";
      return MGPM_ERR_FORMAT;
    }
    max_id = std::max(max_id, std::max(a, b));
    pairs.push_back((uint64_t(a) << 32) | uint64_t(b));
  }
  if (format == 1 && n_decl < 0) {
    err = path + ": no size line";
    return MGPM_ERR_FORMAT;
  }
  if (format == 1 && entries >= 0 && static_cast<int64_t>(pairs.size()) != entries) {
    err = path + ": " + std::to_string(pairs.size()) + " entries, size line says " +
          std::to_string(entries);
    return MGPM_ERR_FORMAT;
  }
  const int64_t n = format == 1 ? n_decl : max_id + 1;
  if (n >= (int64_t(1) << 31)) {
    err = path + ": too many vertices for int32 ids";
    return MGPM_ERR_FORMAT;
  }
  csr_from_draws(static_cast<int32_t>(std::max<int64_t>(n, 0)), pairs, g);
  return MGPM_OK;
}

}  // namespace mgpm
