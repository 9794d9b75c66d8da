// Host-side CSR graph: loader/writer of the reference's on-disk format,
// validation, synthetic generators and the degree-oriented DAG.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace mgpm {

// Undirected simple graph in CSR form: rows sorted ascending, no self-loops,
// no duplicates, symmetric.  nnz = 2m.
struct CsrGraph {
  int32_t n = 0;
  int64_t nnz = 0;
  int32_t meta_max_degree = 0;  // value stored in graph.meta.txt, if loaded
  std::vector<int64_t> rowptr;  // n + 1
  std::vector<int32_t> colidx;  // nnz
  // Optional labels (graph.cc:37-82): one 8-bit label per vertex, one
  // 16-bit label per CSR entry; empty when not loaded or set.
  int32_t num_vertex_classes = 0, num_edge_classes = 0;
  std::vector<uint8_t> vlabels;   // n
  std::vector<uint16_t> elabels;  // nnz
  int32_t degree(int32_t v) const {
    return static_cast<int32_t>(rowptr[v + 1] - rowptr[v]);
  }
  int32_t max_degree() const;
};

// Reference format (graph.cc:18-35): <prefix>.meta.txt holds ten integers
// "n n_edges vid_size eid_size vlabel_size elabel_size max_degree feat_len
// num_vertex_classes num_edge_classes"; <prefix>.vertex.bin holds n+1 int64
// row offsets and <prefix>.edge.bin holds n_edges int32 column ids.
// Returns an MGPM status code and fills `err` on failure.
// flags: kLoadVertexLabels / kLoadEdgeLabels also read
// <prefix>.vlabel.bin (n uint8) / <prefix>.elabel.bin (n_edges uint16),
// checked as graph.cc:37-82 asserts: 0 < num_vertex_classes < 255 with
// exactly that many distinct vertex labels, 0 < num_edge_classes < 65535
// with at most that many distinct edge labels.  A missing label file gets
// labels 1..classes, as the reference's fallback does (graph.cc:49-55,
// 71-77), from a hash of the index instead of rand().
constexpr int kLoadVertexLabels = 1;
constexpr int kLoadEdgeLabels = 2;
int read_graph(const std::string &prefix, CsrGraph &g, std::string &err,
               int flags = 0);
// Writes the labels (and their class counts in the meta file) when set.
int write_graph(const std::string &prefix, const CsrGraph &g, std::string &err);
// Checks labels against their class counts (the rules of read_graph for
// labels read from a file).
int check_labels(const CsrGraph &g, std::string &err, bool check_vertex = true,
                 bool check_edge = true);
// Neighbourhood label frequency of v (Graph::BuildNLF, graph.cc:483-495):
// (label, neighbours of v with that label), ascending by label.
std::vector<std::pair<int32_t, int32_t>> nlf(const CsrGraph &g, int32_t v);
// Every vertex's NLF as a CSR: offsets (n + 1) and, per vertex, its
// distinct neighbour labels ascending with their counts.  Arrays sized by
// a first call with offsets only (labels/counts null) or by offsets[n].
void nlf_csr(const CsrGraph &g, int64_t *offsets, int32_t *labels,
             int32_t *counts);
// Vertices per vertex label, labels 0..num_vertex_classes
// (Graph::computeLabelsFrequency, graph.cc:440-457); false, where the
// reference asserts, if a label exceeds num_vertex_classes.
bool label_frequency(const CsrGraph &g, std::vector<int64_t> &freq);

// Edge files (in place of the reference's closed datasets/converter and
// datasets/tsv2mtx.py): format 1 = MatrixMarket coordinate ("%%MatrixMarket"
// banner, '%' comments, a "rows cols entries" line, then "i j [value]"
// lines), format 2 = an edge list with one "src dst [...]" per line
// (GraphChallenge TSV; '#' and '%' lines are comments).  Vertex ids start at
// `base` (MatrixMarket and GraphChallenge files use 1).  The graph is
// symmetrized; self-loops and duplicates are dropped.  Returns an MGPM
// status (MGPM_ERR_IO, MGPM_ERR_FORMAT) with `err` set.
int read_edge_file(const std::string &path, int format, int base, CsrGraph &g,
                   std::string &err);

// Checks the invariants the counting kernels rely on (the reference assumes
// them without checking): offsets monotone and in range, ids in [0,n), rows
// strictly increasing (sorted, no duplicates), no self-loops, symmetric.
int check_simple(const CsrGraph &g, std::string &err);

// Builds a simple undirected CSR from directed draws packed as (src<<32|dst):
// symmetrizes, drops self-loops and duplicates.  Consumes `pairs`.
void csr_from_draws(int32_t n, std::vector<uint64_t> &pairs, CsrGraph &g);

// Counter-based generators: draw i depends only on (seed, i).
void generate_rmat(int scale, int edge_factor, double a, double b, double c,
                   uint64_t seed, CsrGraph &g);
void generate_er(int32_t n, int64_t draws, uint64_t seed, CsrGraph &g);

// Degree-oriented DAG of graph.cc:196-235: arc s->t iff deg(t) > deg(s) or
// (deg(t) == deg(s) and t > s), deg = undirected degree.
struct Dag {
  std::vector<int64_t> rowptr;   // n + 1
  std::vector<int32_t> colidx;   // m
  std::vector<int32_t> out_deg;  // n
  int32_t max_out = 0;
};
void build_dag(const CsrGraph &g, Dag &d);

// Small/large division of graph.cc:237-312 (used for parity of the
// preprocessing only; the triangle kernels use their own table/stream plan).
struct Division {
  std::vector<int64_t> small_rowptr, large_rowptr;
  std::vector<int32_t> small_colidx, large_colidx;
};
void build_division(const CsrGraph &g, const Dag &d, int32_t deg_thd,
                    Division &out);

// Exclusive prefix sum of `in` into out[0..in.size()] (parallel).
void prefix_sum(const std::vector<int64_t> &in, std::vector<int64_t> &out);

}  // namespace mgpm
