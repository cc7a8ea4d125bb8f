/*
 * kc_oracle.h -- CPU restatement of the PICO k-core methods.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg, and by nothing else.  The product path (pico_amd/, libkcore) never
 * calls it and fails loudly when its HIP build is missing.
 *
 * All graphs are undirected, given as a symmetric CSR: rowptr[V+1] (uint64)
 * and col[rowptr[V]] (uint32), each undirected edge stored in both rows.
 * Core numbers and all bookkeeping are unsigned 32-bit, as in the reference
 * (globals.h:32 `typedef unsigned int uint`).
 */
#ifndef KC_ORACLE_H
#define KC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Batagelj & Zaversnik (2003), "An O(m) Algorithm for Cores Decomposition of
 * Networks": sequential bucket peel.  Returns kmax.  The reference has no CPU
 * k-core; this is the textbook method used as the bit-exact oracle and as the
 * CPU baseline. */
uint32_t orc_bz(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                uint32_t *core);

/* GPP level-synchronous peel (gppeel.cu:12-104 kernels, :184-228 host loop).
 * compat!=0 reproduces the reference's isolated-vertex quirk (level starts at
 * 1, so degree-0 vertices report core 1).  *iterations = the reference's
 * `Number of iterations` (deterministic). */
void orc_gpp(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
             int compat, uint32_t *core, uint64_t *iterations);

/* PeelOne (peelone.cu:14-91, host loop :154-186) under one fixed schedule:
 * parts are visited in vertex order within an iteration and see updates made
 * earlier in the same iteration.  The GPU iteration count is
 * schedule-dependent; the core numbers are not. */
void orc_peelone(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                 uint32_t part_size, uint32_t *core, uint64_t *iterations);

/* Bucket peel of peel_dyn / peelone_dyn (peel-dyn.cc:2-182,
 * peelone-dyn.cc:2-168): per level select {v: deg==level}, drain the queue
 * with clamp-at-level decrements.  *levels = kmax+1 loop trips. */
uint32_t orc_bucket_peel(uint32_t V, const uint64_t *rowptr,
                         const uint32_t *col, uint32_t *core,
                         uint64_t *levels);

/* HistoCore (HistoCore.cu:15-202, host :204-372).  hist_out (optional,
 * rowptr[V]+V words, laid out like the reference's histo array parallel to
 * the degree-prefixed edge list) receives the final histogram. */
void orc_histocore(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                   uint32_t *core, uint64_t *iterations, uint32_t *hist_out);

/* CntCore (CntCore.cu:17-264, host :266-439). */
void orc_cntcore(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                 uint32_t *core, uint64_t *iterations);

/* NbrCore (NbrCore.cu:12-167, host :169-320). */
void orc_nbrcore(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                 uint32_t *core, uint64_t *iterations);

/* Plain Jacobi h-index iteration c <- min(c, H(c)) from c = degree, counting
 * rounds until no change (+1 for the detecting round).  Independent of the
 * three frontier schemes above; pins their iteration counts. */
void orc_hindex_jacobi(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                       uint32_t *core, uint64_t *rounds);

/* Brute force for tiny graphs: for k = 1.. repeatedly delete vertices of
 * degree < k from the k-core candidate set (definition of the k-core). */
void orc_bruteforce(uint32_t V, const uint64_t *rowptr, const uint32_t *col,
                    uint32_t *core);

/* Invariant check: returns 0 when `core` satisfies c(v) <= deg(v), every v
 * has >= c(v) neighbours with c >= c(v), and c(v) = H(c)(v) (the h-index of
 * neighbour cores).  Otherwise returns 1 + the first failing vertex id. */
uint64_t orc_check_invariants(uint32_t V, const uint64_t *rowptr,
                              const uint32_t *col, const uint32_t *core);

#ifdef __cplusplus
}
#endif
#endif
