/*
 * sccgen.h — host-side graph container, CSR/reverse-CSR builder, seeded
 * synthetic generators and text I/O for the SCC engine (libsccgen.so).
 *
 * The CSR layout is the one the reference's graph container exposes:
 *   begin[N+1], node_idx[M]       out-edges, rows sorted by destination
 *   r_begin[N+1], r_node_idx[M]   in-edges (transpose), rows sorted by source
 * (reference: gm_graph/inc/gm_graph.h:154-160; rows are semi-sorted after
 *  freeze, gm_graph/src/gm_graph.cc:99-100, and the reverse rows too, :276).
 *
 * All ids and offsets are int32 (reference: gm_graph/inc/gm_graph_typedef.h:17-18),
 * so M must stay below 2^31.  Every generator is counter-based (SplitMix64
 * keyed by (seed, stream, index)), so its output is independent of the number
 * of host threads used.
 *
 * Error convention: functions returning int give 0 on success and a negative
 * code on failure; sccgen_last_error() returns a message for the calling thread.
 */
#ifndef SCCGEN_H
#define SCCGEN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sccgen_graph sccgen_graph; /* opaque: N, M and the four CSR arrays */

const char* sccgen_last_error(void);

/* ---- building from an edge list --------------------------------------- */

/* Build CSR + reverse CSR from n vertices and m directed edges (src[i]->dst[i]).
 * dedup=0 keeps self-loops and duplicate edges, as the reference loader does
 *   (reference: src/common_main.h:61-83 reads edges; gm_graph.cc:459-481 adds them
 *    unchanged; make_reverse_edges gm_graph.cc:193-287 builds the transpose).
 * dedup=1 removes self-loops and duplicates (the synthetic-workload convention).
 * Returns NULL on error. */
sccgen_graph* sccgen_from_edges(int32_t n, int64_t m, const int32_t* src,
                                const int32_t* dst, int dedup);

/* Read the reference's text edge list: one "u v" pair per line, 1-indexed,
 * single-space separated, N = max id seen (reference: src/common_main.h:61-83).
 * Duplicates and self-loops are kept. */
sccgen_graph* sccgen_read_edgelist(const char* path);

/* Write a graph as the reference's text edge list (1-indexed "u v" lines in CSR
 * order).  Returns 0 on success. */
int sccgen_write_edgelist(const sccgen_graph* g, const char* path);

/* gm binary CSR (reference: gm_graph/src/gm_graph_binary_loader.cc:21-45,
 * load :48-205, store :207-248): big-endian words
 *   magic 0x03939999 | sizeof(node_t)=4 | sizeof(edge_t)=4 | N | M |
 *   begin[N+1] | node_idx[M].
 * The reader also takes the older little-endian ("flipped") files the
 * reference accepts, rejects 8-byte ids (the reference's int32 build does too),
 * sorts each row and builds the reverse CSR, as load_binary does.  Returns NULL
 * on error (sccgen_last_error says why). */
sccgen_graph* sccgen_read_gm_binary(const char* path);
/* Write the forward CSR in that format (big-endian).  Returns 0 on success. */
int sccgen_write_gm_binary(const sccgen_graph* g, const char* path);

/* ---- seeded synthetic generators -------------------------------------- */

/* Directed R-MAT: n = 2^scale vertices, round(edge_factor * n) sampled edges
 * with quadrant probabilities (a, b, c, 1-a-b-c); if permute, vertex ids are
 * relabelled by a seeded Fisher-Yates permutation; self-loops and duplicates
 * are removed when dedup.  If n_keep > 0 (and n_keep <= 2^scale), edges with an
 * endpoint >= n_keep after permutation are rejected and the graph has n_keep
 * vertices (used for non-power-of-two sizes). */
sccgen_graph* sccgen_rmat(int scale, double edge_factor, double a, double b,
                          double c, uint64_t seed, int permute, int dedup,
                          int64_t n_keep);

/* Planted-SCC graph.  One giant SCC of giant_frac*n vertices and n_medium SCCs
 * with sizes drawn from a power law s^-alpha on [size_min, size_max]; each
 * planted SCC is a random Hamiltonian cycle plus intra_out random intra-SCC
 * out-edges per vertex.  Remaining vertices become 2-cycles (fraction
 * pair_frac of them) and singletons.  Inter-component edges go from lower to
 * higher random topological rank, inter_per_vertex per vertex on average.
 * Vertex ids are permuted by a seeded permutation; self-loops/duplicates are
 * removed.  If comp_out != NULL it receives, per vertex, the planted component
 * index (length n), which fixes the true SCC partition. */
sccgen_graph* sccgen_planted(int32_t n, double giant_frac, int32_t n_medium,
                             int32_t size_min, int32_t size_max, double alpha,
                             int32_t intra_out, double inter_per_vertex,
                             double pair_frac, uint64_t seed, int32_t* comp_out);

/* ---- accessors ---------------------------------------------------------- */

int32_t sccgen_num_nodes(const sccgen_graph* g);
int64_t sccgen_num_edges(const sccgen_graph* g);
/* Copy the CSR arrays out; any pointer may be NULL to skip that array. */
void sccgen_copy(const sccgen_graph* g, int32_t* begin, int32_t* node_idx,
                 int32_t* r_begin, int32_t* r_node_idx);
/* Borrow a pointer to one array (0 begin, 1 node_idx, 2 r_begin, 3 r_node_idx);
 * valid until sccgen_free. */
const int32_t* sccgen_array(const sccgen_graph* g, int which);
void sccgen_free(sccgen_graph* g);

/* ---- dynamic update stream --------------------------------------------- */

/* A seeded stream of edge-update batches against a base graph.  Batch sizes
 * are floor(exp(U[ln lo, ln hi])); each batch is half inserts of uniform
 * random vertex pairs (u != v, not already present) and half deletes chosen
 * uniformly among the edges present at that point of the stream. */
typedef struct sccgen_updates sccgen_updates;
sccgen_updates* sccgen_updates_create(const sccgen_graph* base, uint64_t seed,
                                      int64_t size_lo, int64_t size_hi);
/* Size of the next batch (inserts + deletes) without drawing it. */
int64_t sccgen_updates_peek(const sccgen_updates* u);
/* Sizes of the following batches: log-uniform in [size_lo, size_hi].  0 on success. */
int sccgen_updates_set_sizes(sccgen_updates* u, int64_t size_lo, int64_t size_hi);
/* Draw the next batch: fills up to cap entries of each list; returns 0 and the
 * counts, or <0 if cap is too small. */
int sccgen_updates_next(sccgen_updates* u, int64_t cap, int32_t* ins_src,
                        int32_t* ins_dst, int64_t* n_ins, int32_t* del_src,
                        int32_t* del_dst, int64_t* n_del);
/* Current edge count of the tracked graph. */
int64_t sccgen_updates_num_edges(const sccgen_updates* u);
/* Materialise the tracked current graph (caller frees). */
sccgen_graph* sccgen_updates_graph(const sccgen_updates* u);
void sccgen_updates_free(sccgen_updates* u);

/* ---- helpers ------------------------------------------------------------ */

/* Canonicalise labels: out[v] = min vertex id among vertices sharing in[v].
 * Labels must be in [0, n). */
int sccgen_canonicalize(int32_t n, const int32_t* labels, int32_t* out);

/* SplitMix64-based counter RNG used by every generator: value for (key, i). */
uint64_t sccgen_rng(uint64_t key, uint64_t i);

#ifdef __cplusplus
}
#endif
#endif /* SCCGEN_H */
