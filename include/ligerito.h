/*
 * ligerito.h — C ABI of the MI355X (gfx950) Ligerito prover hot path.
 *
 * Drop-in boundary for the public API of the reference crate
 * `cryptoutils` (ashutosh1206/ligerito-rust).  Each entry point names the
 * reference item it replaces (file:line).  Conventions:
 *
 *  - Field elements cross the ABI as little-endian bytes of field_bits/8
 *    bytes each (BinaryField::to_bytes, binary_field.rs:161-163):
 *    uint32_t arrays for GF(2^32), pairs of uint64_t (low word first) for
 *    GF(2^128).  field_bits is 32 or 128.
 *  - Polynomials and encoded matrices use the reference layout: a ligero
 *    matrix is column-major, column j = poly[j*m .. (j+1)*m] (ligero.rs:11-27).
 *  - Every function returns LIG_OK (0) or a negative LIG_E_* status; the
 *    reference's panics (assert!, unwrap, expect) map to LIG_E_ASSERT and
 *    nothing unwinds across the ABI.  lig_last_error() describes the last
 *    failure on the calling thread.
 *  - Buffers are caller-owned; handles are library-owned and freed with the
 *    matching *_free.  Calls on one lig_ctx are stream-ordered and must not
 *    overlap from several threads.
 */
#ifndef LIGERITO_H
#define LIGERITO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LIG_OK 0
#define LIG_E_INVALID (-1)    /* bad argument (a reference assert on inputs) */
#define LIG_E_ASSERT (-2)     /* a reference assert_eq!/unwrap would panic */
#define LIG_E_HIP (-3)        /* HIP runtime failure */
#define LIG_E_OOM (-4)        /* device allocation failed */
#define LIG_E_NO_DEVICE (-5)  /* no usable GPU */
#define LIG_E_MALFORMED (-6)  /* undecodable proof bytes */
#define LIG_E_CAPACITY (-7)   /* caller buffer too small */
#define LIG_E_MERKLE (-8)     /* verifier: Merkle opening check failed (verifier.rs:50,97-106,135-144) */
#define LIG_E_LIGERO (-9)     /* verifier: ligero_verify row check failed (ligero.rs:124) */

/* Description of the last error on this thread ("" if none). */
const char *lig_last_error(void);
/* Number of visible GPUs (0 on a host without one). */
int lig_device_count(void);

/* ------------------------------------------------------------ context -- */
/* Device + stream + plan caches.  Replaces the reference's implicit global
 * rayon pool (no counterpart type in the reference). */
typedef struct lig_ctx lig_ctx;
int lig_ctx_create(int device, lig_ctx **out);
void lig_ctx_destroy(lig_ctx *ctx);
/* Instrumentation (no counterpart in the reference; for benchmarks): with
 * profiling on, the library records a HIP event pair on its stream around
 * each phase of its work and accumulates the elapsed time per phase.
 * Phases: "encode0" (Reed–Solomon encode of the first commit of a prove:
 * transpose+scale and FFT pass kernels), "encode" (encodes of the other
 * commits), "merkle" (leaf hashing + tree levels), "fold", "induce",
 * "sumcheck", "open", and "pass" (each FFT pass launch of the first commit
 * of a prove on its own).  Enabling (or disabling) resets the totals. */
int lig_ctx_profile(lig_ctx *ctx, int enable);
int lig_ctx_profile_read(lig_ctx *ctx, const char *phase, double *total_ms, uint64_t *count);
/* Microbenchmarks for the rooflines (no counterpart in the reference):
 * which = 0: GF(2^32) products per second by wave-uniform constants with
 * the FFT passes' nibble tables in LDS; 1: general GF(2^32) products per
 * second (carry-less product from integer multiplies, as the induce and
 * sumcheck kernels multiply); 2: 32-bit integer VALU instructions per
 * second summed over lanes; 3: GF(2^32) products per second by constants
 * with 5-bit-window tables (seven 4-byte LDS reads per product instead of
 * eight); 4: LDS bytes read per second with the GF(2^128) table pattern
 * (16-byte reads, 16 distinct entries per wave instruction); 5: GF(2^32)
 * products per second by a constant from its 32 multiples in registers
 * (bit masks, no tables); 6: GF(2^8) products per second by a constant
 * through log/exp tables in LDS (the subfield step of a tower
 * representation).  Best of five launches on the context's GPU. */
int lig_bench_rate(lig_ctx *ctx, int which, double *rate);
/* The hipStream_t all work of this context is ordered on. */
void *lig_ctx_stream(lig_ctx *ctx);
int lig_ctx_synchronize(lig_ctx *ctx);

/* -------------------------------------------------------- Reed–Solomon -- */
/* ReedSolomonEncoding<F> (reed_solomon.rs:7-17). */
typedef struct lig_rs lig_rs;
/* ReedSolomonEncoding::new (reed_solomon.rs:49-64): twiddles for the block
 * length and pis for the message length; both powers of two,
 * message_length < block_length.  Device tables are built here (setup). */
int lig_rs_new(lig_ctx *ctx, int field_bits, size_t message_length, size_t block_length, lig_rs **out);
void lig_rs_free(lig_rs *rs);
/* message_length() / block_length() (reed_solomon.rs:66-72). */
size_t lig_rs_message_length(const lig_rs *rs);
size_t lig_rs_block_length(const lig_rs *rs);
/* encode (reed_solomon.rs:90-102): systematic; host buffers, out has
 * block_length elements. */
int lig_rs_encode(lig_rs *rs, const void *message, void *out);
/* encode_non_systematic (reed_solomon.rs:104-113): in place over a host
 * buffer of block_length elements. */
int lig_rs_encode_non_systematic(lig_rs *rs, void *message);
/* Device variant of encode_non_systematic over n_cols columns stored
 * column-major (each column block_length elements), as encode_cols
 * (ligero.rs:29-38) applies it.  n_cols must be a power of two
 * (LIG_E_INVALID otherwise).  Synchronous. */
int lig_rs_encode_non_systematic_cols_dev(lig_rs *rs, void *d_columns, size_t n_cols);
/* No counterpart in the reference: the same encode for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc, hipFree or
 * allocating routine: its row-major working copy lives in a static 64 MiB
 * device array, one per GPU, held from the library's load to the end of the
 * process.  block_length * n_cols * field_bits/8 must not exceed 64 MiB
 * (LIG_E_INVALID otherwise).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (from several threads or contexts) race.
 * Synchronous; results equal lig_rs_encode_non_systematic_cols_dev's. */
int lig_rs_encode_non_systematic_cols_dev_noalloc(lig_rs *rs, void *d_columns, size_t n_cols);
/* Autotuning of the encode (no counterpart in the reference; no effect on
 * any result): the FFT passes of an encode of n_cols columns can be launched
 * in several shapes — kmax, the most butterfly stages per pass (GF(2^32)
 * 2..6, GF(2^128) 1..3), and group, the most tiles (GF(2^32), up to 8) or
 * 64-column jobs (GF(2^128), up to 16) per workgroup.  The first commit of
 * n_cols columns on a plan times every candidate on this GPU and keeps the
 * fastest (not for small matrices, nor with LIG_AUTOTUNE=0 in the
 * environment; 0 then stands for the built-in default).  lig_rs_autotune
 * tunes now if that has not happened and returns the choice;
 * lig_rs_set_tuning fixes it (0 = default). */
int lig_rs_autotune(lig_rs *rs, size_t n_cols, int *kmax, int *group);
int lig_rs_set_tuning(lig_rs *rs, size_t n_cols, int kmax, int group);
/* compute_twiddles output of this plan (binary_fft.rs:30-58): block_length-1
 * elements. */
int lig_rs_twiddles(const lig_rs *rs, void *out);

/* ------------------------------------------------------------- Merkle -- */
/* build_merkle_tree (merkle_tree.rs:22-60) over n_leaves fixed-length
 * leaves (host buffer, leaf i = leaves[i*leaf_len ..]).  layers_out receives
 * every layer, leaf hashes first, 32*(2*n_leaves-1) bytes. */
int lig_build_merkle_tree(lig_ctx *ctx, const void *leaves, size_t n_leaves, size_t leaf_len, uint8_t *layers_out);
/* build_merkle_tree over leaves of any lengths (its AsRef<[u8]> leaves,
 * merkle_tree.rs:22-60): leaf i is bytes[offsets[i] .. offsets[i+1]), so
 * offsets holds n_leaves + 1 non-decreasing byte offsets; leaves may be empty
 * and at most 256 KiB each.  n_leaves must be a power of two (LIG_E_ASSERT,
 * merkle_tree.rs:26).  layers_out as for lig_build_merkle_tree. */
int lig_build_merkle_tree_var(lig_ctx *ctx, const void *bytes, const size_t *offsets, size_t n_leaves,
                              uint8_t *layers_out);
/* verify (merkle_tree.rs:164-188) with leaves of any lengths, laid out as
 * for lig_build_merkle_tree_var (n_queries + 1 offsets): 1 (valid) or 0. */
int lig_merkle_verify_var(const uint8_t root[32], const uint8_t *proof, size_t n_nodes, size_t depth, const void *bytes,
                          const size_t *offsets, const size_t *indices, size_t n_queries);
/* prove (merkle_tree.rs:98-111) from the layers written above; queries
 * sorted and distinct.  Writes up to cap_nodes 32-byte nodes. */
int lig_merkle_prove(const uint8_t *layers, size_t n_leaves, const size_t *queries, size_t n_queries, uint8_t *proof_out,
                     size_t cap_nodes, size_t *n_nodes);
/* verify (merkle_tree.rs:164-188): returns 1 (valid) or 0. */
int lig_merkle_verify(const uint8_t root[32], const uint8_t *proof, size_t n_nodes, size_t depth, const void *leaves,
                      size_t leaf_len, const size_t *indices, size_t n_queries);

/* ---------------------------------------------------------- transcript -- */
/* FS (emulated_fs.rs:6-75). */
typedef struct lig_fs lig_fs;
int lig_fs_new(uint32_t seed, lig_fs **out);             /* FS::new (:12-16) */
void lig_fs_free(lig_fs *fs);
int lig_fs_absorb(lig_fs *fs, const void *bytes, size_t len); /* absorb (:18-20) */
int lig_fs_squeeze(lig_fs *fs, uint8_t out16[16]);      /* squeeze (:22-32), u128 LE */
int lig_fs_get_field(lig_fs *fs, int field_bits, void *out); /* get_field (:34-54) */
size_t lig_fs_get_query(lig_fs *fs, size_t n);          /* get_query (:56-59) */
/* get_distinct_queries (:61-75); s <= n is required (the reference loops
 * forever otherwise): LIG_E_INVALID. */
int lig_fs_get_distinct_queries(lig_fs *fs, size_t n, size_t s, size_t *out);

/* --------------------------------------------------------- ligero_commit -- */
/* RecursiveLigeroWitness<F> (data_structures.rs:32-40), device-resident. */
typedef struct lig_witness lig_witness;
/* ligero_commit (ligero.rs:70-95): poly has m*n elements (host), rs must
 * have message_length m and block_length 4m.  n must be a power of two
 * (every configuration of the reference uses one); other n give
 * LIG_E_INVALID. */
int lig_ligero_commit(lig_rs *rs, const void *poly, size_t m, size_t n, lig_witness **out);
/* Same with poly already in device memory. */
int lig_ligero_commit_dev(lig_rs *rs, const void *d_poly, size_t m, size_t n, lig_witness **out);
int lig_witness_root(const lig_witness *w, uint8_t root[32]);    /* tree.last() */
size_t lig_witness_num_rows(const lig_witness *w);
size_t lig_witness_num_cols(const lig_witness *w);
/* flat_mat in the reference's column-major layout (4m*n elements). */
int lig_witness_flat_mat(const lig_witness *w, void *out);
/* every Merkle layer, leaf hashes first (32*(2*num_rows-1) bytes). */
int lig_witness_layers(const lig_witness *w, uint8_t *out);
/* extract_row (ligero.rs:40-51). */
int lig_witness_extract_row(const lig_witness *w, size_t row, void *out);
void lig_witness_free(lig_witness *w);

/* ---------------------------------------------------------------- config -- */
/* ProverConfig<F> (data_structures.rs:9-21).  log_n = 20 and 24 build
 * exactly hardcoded_config_20 / hardcoded_config_24 (config.rs:6-47,69-112);
 * other log_n in [16, 27] follow the same shape rules. */
typedef struct lig_prover_config lig_prover_config;
int lig_prover_config_new(lig_ctx *ctx, int field_bits, int log_n, lig_prover_config **out);
/* ProverConfig built from its fields (data_structures.rs:9-21), as
 * hardcoded_config_20/24 fill them (config.rs:19-46): initial_dims =
 * (rows, cols) of the first matrix, dims[2i], dims[2i+1] = (rows, cols) of
 * recursive step i, ks[i] its sumcheck rounds.  The shape must satisfy what
 * prover() and verifier() rely on, else LIG_E_INVALID names the broken rule:
 * recursive_steps >= 1 (prover.rs:198); initial_k = log2 initial_dims[1] and
 * ks[i] = log2 dims[i].cols, both >= 1 (prover.rs:50-53,97-102);
 * initial_dims[0] = dims[0].rows * dims[0].cols and dims[i].rows =
 * dims[i+1].rows * dims[i+1].cols; 4 * rows >= 148 for every matrix
 * (emulated_fs.rs:65 never ends otherwise); powers of two
 * (reed_solomon.rs:49-64).  Limits of this library: at most 8 steps and
 * 2^27 coefficients, rows of at most 4 KiB, a final message of at most
 * 2 MiB.  The input of lig_prover is initial_dims[0] * initial_dims[1]
 * elements. */
int lig_prover_config_from_dims(lig_ctx *ctx, int field_bits, size_t recursive_steps, const size_t initial_dims[2],
                                size_t initial_k, const size_t *dims, const size_t *ks, lig_prover_config **out);
/* The generated configuration for 2^log_n coefficients (the shape
 * lig_prover_config_new builds; config.rs:6-47,69-112 at log_n = 20, 24),
 * in the layout of lig_prover_config_from_dims; no device needed.  dims and
 * ks may be NULL (only the first three outputs are written); otherwise they
 * hold cap_steps steps (LIG_E_CAPACITY if more are needed). */
int lig_prover_config_shape(int log_n, size_t *recursive_steps, size_t initial_dims[2], size_t *initial_k,
                            size_t *dims, size_t *ks, size_t cap_steps);
/* The shape of a configuration, as lig_prover_config_shape writes it. */
int lig_prover_config_dims(const lig_prover_config *cfg, size_t *recursive_steps, size_t initial_dims[2],
                           size_t *initial_k, size_t *dims, size_t *ks, size_t cap_steps);
void lig_prover_config_free(lig_prover_config *cfg);

/* ---------------------------------------------------------------- prove -- */
/* FinalizedLigeritoProof<F> (data_structures.rs:91-103). */
typedef struct lig_proof lig_proof;
/* prover (prover.rs:15-199): poly has 2^log_n elements (host).  fs_seed is
 * the transcript seed; the reference hard-codes 1234 (prover.rs:29). */
int lig_prover(lig_prover_config *cfg, const void *poly, uint32_t fs_seed, lig_proof **out);
/* Same with poly already in device memory (left unmodified). */
int lig_prover_dev(lig_prover_config *cfg, const void *d_poly, uint32_t fs_seed, lig_proof **out);
/* Canonical encoding (the reference has none), little-endian in struct
 * field order: "LGRT" | u32 version 1 | u32 field_bytes | initial root[32] |
 * RLP(initial) | u64 n + n roots[32] | u64 n + n RLPs | u64 len + yr |
 * RLP(final) | u64 n + n (s0,s1,s2) triples, where RLP = u64 n_rows |
 * u64 n_cols | rows (row-major) | u64 n_nodes | nodes[32].  With buf == NULL
 * only *len is set. */
int lig_proof_serialize(const lig_proof *p, uint8_t *buf, size_t cap, size_t *len);
int lig_proof_deserialize(int field_bits, const uint8_t *buf, size_t len, lig_proof **out);
/* ProofSize::proof_size (data_structures.rs:105-194). */
size_t lig_proof_size(const lig_proof *p);
void lig_proof_free(lig_proof *p);

/* --------------------------------------------------------------- verify -- */
/* verifier (verifier.rs:10-164) on the configuration for 2^log_n (host
 * CPU).  Returns 1 (accept), 0 (final sumcheck check false), or
 * LIG_E_MERKLE / LIG_E_LIGERO / LIG_E_ASSERT where the reference panics. */
int lig_verifier(int field_bits, int log_n, const lig_proof *p, uint32_t fs_seed);
/* The same verifier with its basis inductions (induce_sumcheck_poly at
 * verifier.rs:54-61,148-155, the bulk of its arithmetic) on the context's
 * GPU; same results and status codes as lig_verifier. */
int lig_verifier_dev(lig_ctx *ctx, int field_bits, int log_n, const lig_proof *p, uint32_t fs_seed);
/* verifier (verifier.rs:10-164) on an explicit VerifierConfig
 * (data_structures.rs:23-30): initial_dim = log2 of the first matrix's rows,
 * log_dims[i] = log2 of step i's rows, initial_k and ks[i] the fold sizes
 * (config.rs:49-67,114-134 for the two hardcoded ones).  A shape that
 * breaks the rules of lig_prover_config_from_dims gives LIG_E_INVALID. */
int lig_verifier_with_config(int field_bits, size_t recursive_steps, size_t initial_dim, const size_t *log_dims,
                             size_t initial_k, const size_t *ks, const lig_proof *p, uint32_t fs_seed);
/* The same with the basis inductions on the context's GPU. */
int lig_verifier_with_config_dev(lig_ctx *ctx, int field_bits, size_t recursive_steps, size_t initial_dim,
                                 const size_t *log_dims, size_t initial_k, const size_t *ks, const lig_proof *p,
                                 uint32_t fs_seed);

/* ---------------------------------------------------------------- batch -- */
/* No counterpart in the reference (its prover runs one proof at a time):
 * many independent proves on one device.  A batch holds n_streams workers,
 * each with its own HIP stream, device memory pool, pinned staging and prover
 * configurations (built on first use of a size), so the proves of different
 * jobs overlap on the GPU.  GF(2^32) jobs of one size are proved in lockstep
 * in gangs of up to four (every kernel launched once for the gang);
 * LIG_BATCH_GANG=0 in the environment proves each job on its own. */
typedef struct lig_batch lig_batch;
int lig_batch_create(int device, int field_bits, int n_streams, lig_batch **out);
void lig_batch_destroy(lig_batch *b);
/* Proves n_jobs polynomials: job i has 2^log_n[i] elements (log_n in
 * [16, 27]) at polys[i] — host memory, or device memory when
 * polys_on_device — and transcript seed fs_seeds[i].  out[i] receives a proof
 * identical to lig_prover's for the same input and seed.  Returns when every
 * job is done; on error no proof is returned (out[] all NULL). */
int lig_prove_batch(lig_batch *b, size_t n_jobs, const int *log_n, const void *const *polys, int polys_on_device,
                    const uint32_t *fs_seeds, lig_proof **out);
/* -------------------------------------------------------------- serving -- */
/* No counterpart in the reference: a prove server with continuous batching.
 * n_slots workers (as in a batch) take requests first come, first served; a
 * request is admitted when a worker is free and its device-memory estimate
 * fits in hbm_budget bytes beside the requests already running (a request
 * larger than the whole budget runs alone).  Requests that arrive while
 * others run are admitted as slots and memory free up, and their GPU work
 * overlaps. */
typedef struct lig_server lig_server;
int lig_serve_start(int device, int field_bits, int n_slots, uint64_t hbm_budget, lig_server **out);
/* Queues a prove of 2^log_n elements at poly (host memory; it must stay
 * valid until the request completes) with transcript seed fs_seed. */
int lig_serve_submit(lig_server *s, int log_n, const void *poly, uint32_t fs_seed, uint64_t *ticket);
/* 1 with *out set (and the submit-to-completion latency in ms) once the
 * request is done — the ticket is then released — 0 while it is pending, or
 * the request's negative status if it failed. */
int lig_serve_poll(lig_server *s, uint64_t ticket, lig_proof **out, double *latency_ms);
/* As lig_serve_poll, waiting for completion. */
int lig_serve_wait(lig_server *s, uint64_t ticket, lig_proof **out, double *latency_ms);
/* Prepares every slot for requests of the count sizes in log_n: builds its
 * prover configuration and caches the device blocks one prove of each size
 * takes (one prove of a zero polynomial per slot and size), so that the
 * first requests of those sizes pay for neither.  Waits for the running
 * requests to finish and admits no queued request until it returns; stops
 * at the first slot and size whose memory would exceed hbm_budget.  Sets
 * *warmed (may be NULL) to the number of (slot, size) pairs prepared.  The
 * zero input is a pinned host buffer of the largest size, held for the
 * call.  A second call while one runs fails with LIG_E_INVALID; a call
 * must not overlap lig_serve_stop.
 * No counterpart in the reference (it has no server). */
int lig_serve_warm(lig_server *s, const int *log_n, size_t count, size_t *warmed);
/* Finishes the running requests, fails the queued ones, frees the server. */
void lig_serve_stop(lig_server *s);
#ifdef __cplusplus
}
#endif
#endif /* LIGERITO_H */
