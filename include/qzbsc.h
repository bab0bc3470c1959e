/*
 * qzbsc — MI355X (gfx950) block-sorting engine for qz's BSC stream path.
 *
 * Part 1 (libbsc drop-in): the four symbols qz-lib binds through FFI, with
 * libbsc's signatures, error codes and buffer contract. A build of this
 * library named liblibbsc.a / liblibbsc.so links in place of libbsc.
 *
 * Part 2 (BWT stages): libbsc's bwt.h entry points, plus batch and
 * device-resident forms used by the batched and serving paths.
 *
 * Error codes (crates/qz-lib/src/compression/bsc.rs:18-24):
 *   0 no error, -1 bad parameter, -2 not enough memory, -3 not compressible,
 *   -4 not supported, -5 unexpected end of block, -6 data corrupt;
 *   -7 GPU runtime error (this library only).
 */
#ifndef QZBSC_H
#define QZBSC_H

#ifdef __cplusplus
extern "C" {
#endif

#define QZBSC_API __attribute__((visibility("default")))

#define LIBBSC_NO_ERROR 0
#define LIBBSC_BAD_PARAMETER -1
#define LIBBSC_NOT_ENOUGH_MEMORY -2
#define LIBBSC_NOT_COMPRESSIBLE -3
#define LIBBSC_NOT_SUPPORTED -4
#define LIBBSC_UNEXPECTED_EOB -5
#define LIBBSC_DATA_CORRUPT -6
#define QZBSC_GPU_ERROR -7

#define LIBBSC_BLOCKSORTER_BWT 1
#define LIBBSC_CODER_QLFC_STATIC 1
#define LIBBSC_CODER_QLFC_ADAPTIVE 2
#define LIBBSC_CODER_QLFC_FAST 3
/* Coder id this library writes into bits 5-7 of a block's mode word: its own
 * adaptive rank/run model, whose bitstream is not libbsc's QLFC. libbsc
 * defines ids 1-3 only, so upstream libbsc rejects these blocks instead of
 * misreading them, and this library reports libbsc-coded blocks (ids 1-3) as
 * LIBBSC_NOT_SUPPORTED. */
#define QZBSC_CODER_QZ_ADAPTIVE 4
#define LIBBSC_FEATURE_FASTMODE 1
#define LIBBSC_FEATURE_MULTITHREADING 2
#define LIBBSC_FEATURE_CUDA 8 /* "use the GPU" — selects the HIP path here */
#define LIBBSC_HEADER_SIZE 28

/* ---- Part 1: libbsc drop-in ------------------------------------------- */

/* Replaces libbsc bsc_init; bound at crates/qz-lib/src/compression/bsc.rs:94,
 * called once through std::sync::Once (bsc.rs:128-141). */
QZBSC_API int bsc_init(int features);

/* Replaces libbsc bsc_compress; bound at bsc.rs:97-106, called at
 * bsc.rs:172-183 and bsc.rs:234-247. `output` holds n + 28 + 1024 bytes
 * (bsc.rs:170). Returns the block size or a negative error. Parameter sets
 * qz uses: (0, 0, BWT, ADAPTIVE) default, (0, 0, BWT, STATIC), (16, 128, BWT,
 * ADAPTIVE) with LZP. Every coder request (static, adaptive, fast) is served
 * by this library's adaptive model and recorded in the block header as
 * QZBSC_CODER_QZ_ADAPTIVE; LZP blocks use this library's token format. */
QZBSC_API int bsc_compress(const unsigned char* input, unsigned char* output, int n, int lzpHashSize,
                           int lzpMinLen, int blockSorter, int coder, int features);

/* Replaces libbsc bsc_block_info; bound at bsc.rs:109-115, called at
 * bsc.rs:401-409 with the whole block as the header. */
QZBSC_API int bsc_block_info(const unsigned char* blockHeader, int headerSize, int* pBlockSize, int* pDataSize,
                             int features);

/* Replaces libbsc bsc_decompress; bound at bsc.rs:118-124, called at
 * bsc.rs:421-429 with outputSize == dataSize. Returns 0 on success, and
 * LIBBSC_NOT_SUPPORTED for blocks coded by libbsc's own QLFC coders. */
QZBSC_API int bsc_decompress(const unsigned char* input, int inputSize, unsigned char* output, int outputSize,
                             int features);

/* Batch forms of bsc_compress / bsc_decompress over the same block format.
 * Blocks lie end to end: block b of bsc_compress_batch starts at input +
 * sum of the earlier sizes and writes at most sizes[b] + 28 + 1024 bytes at
 * output + sum over c < b of (sizes[c] + 28 + 1024), its size (or a negative
 * code) in out_sizes[b]; block b of qzbsc_decompress_batch (in_sizes[b]
 * bytes) decodes into output + sum of the earlier out_sizes, out_sizes[b]
 * >= its data size, its code in rcs[b]. One device batch (up to 1 GiB)
 * serves all blocks; the host coder runs on one thread per block. Both
 * return 0 or the first block's error. No counterpart in the reference:
 * they replace qz's rayon fan-out of per-block calls over a stream's
 * blocks (mod.rs:388-413 compress_stream_to_bsc_blocks; the batches of 8
 * blocks of decompress_impl.rs:97-121). */
QZBSC_API int qzbsc_compress_batch(const unsigned char* input, const int* sizes, int nblocks, unsigned char* output,
                                   int* out_sizes, int lzpHashSize, int lzpMinLen, int blockSorter, int coder,
                                   int features);
QZBSC_API int qzbsc_decompress_batch(const unsigned char* input, const int* in_sizes, int nblocks,
                                     unsigned char* output, const int* out_sizes, int features, int* rcs);

/* FASTQ text <-> the header / sequence / quality streams of a .qz archive
 * (host). qzbsc_fastq_to_streams first with hdr, seq, qual NULL: info[0..5]
 * = records, constant sequence length, constant quality length (0 =
 * variable), and the three stream sizes; then with buffers of those sizes
 * and info unchanged, to write them. Replaces records_to_streams
 * (crates/qz-lib/src/compression/mod.rs:280-320, encoding 0, lossless
 * qualities) and detect_const_lengths (mod.rs:329-354). -1 for malformed
 * FASTQ (a record needs four lines, '@' and '+').
 * qzbsc_streams_to_fastq returns the FASTQ size (out may be NULL to size it)
 * or -5 when a stream ends early. Replaces the offset scans and record
 * assembly of decompress_impl.rs:444-636. */
QZBSC_API long long qzbsc_fastq_to_streams(const unsigned char* fastq, long long size, unsigned char* hdr,
                                           unsigned char* seq, unsigned char* qual, long long* info);
QZBSC_API long long qzbsc_streams_to_fastq(const unsigned char* hs, long long hlen, const unsigned char* ss,
                                           long long slen, const unsigned char* qs, long long qlen, long long n,
                                           long long const_seq, long long const_qual, unsigned char* out,
                                           long long cap);

/* ---- Part 2: BWT stages ------------------------------------------------ */

/* libbsc bwt.h bsc_bwt_encode (called inside libbsc's bsc_compress, which
 * qz reaches at bsc.rs:172-183): in-place forward BWT; returns the primary
 * index and writes (n-1)/r trailer indexes, r the largest power of two
 * <= n/8. */
QZBSC_API int qzbsc_bwt_encode(unsigned char* T, int n, unsigned char* num_indexes, int* indexes, int features);

/* libbsc bwt.h bsc_bwt_decode (inside bsc_decompress, bsc.rs:421-429):
 * in-place inverse BWT. */
QZBSC_API int qzbsc_bwt_decode(unsigned char* T, int n, int index, unsigned char num_indexes, int* indexes,
                               int features);

/* Batch forms of the two calls above: nblocks blocks of sizes[b] bytes laid
 * end to end in `in`/`out` (host memory), up to 4 GB in all and each block
 * below 2^31 bytes (LIBBSC_BAD_PARAMETER otherwise). The forward sorts the
 * batch by alphabet class in chunks of whole blocks, the inverse in
 * sub-batches of whole blocks below 2^31 bytes.
 * primary[b] receives the block's primary index; aux (nblocks x 256 ints,
 * may be NULL) receives I[k] = 1 + rank of suffix k*r, k = 0..(n-1)/r.
 * No counterpart in the reference: they replace qz's rayon fan-out of
 * per-block calls (bsc.rs:282-288, decompress_impl.rs:117-120). */
QZBSC_API int qzbsc_bwt_encode_batch(const unsigned char* in, unsigned char* out, const int* sizes, int nblocks,
                                     int* primary, int* aux);
QZBSC_API int qzbsc_bwt_decode_batch(const unsigned char* in, unsigned char* out, const int* sizes, int nblocks,
                                     const int* primary);

/* Device-resident batch forms: d_in / d_out are device pointers. gpu_ms
 * (may be NULL) receives the device time of the transform measured with
 * HIP events on the engine's stream. No counterpart in the reference. */
QZBSC_API int qzbsc_dev_bwt_encode_batch(const void* d_in, void* d_out, const int* sizes, int nblocks,
                                         int* primary, int* aux, float* gpu_ms);
QZBSC_API int qzbsc_dev_bwt_decode_batch(const void* d_in, void* d_out, const int* sizes, int nblocks,
                                         const int* primary, float* gpu_ms);
/* qzbsc_dev_bwt_decode_batch without any allocation, for callers that
 * cannot allocate device memory (code under graph capture, tight serving
 * loops). No counterpart in the reference. The workspace is a 64 MiB array
 * every device the library runs on holds for the life of the process; a
 * batch may hold at most 8 MiB (8,388,608 bytes) in at most 64 blocks, and
 * larger batches return LIBBSC_NOT_ENOUGH_MEMORY. Runs on `stream` (a
 * hipStream_t; NULL for the null stream) and returns after it finishes.
 * Calls on one device share its array and are serialised in this process. */
QZBSC_API int qzbsc_dev_bwt_decode_batch_noalloc(const void* d_in, void* d_out, const int* sizes, int nblocks,
                                                 const int* primary, void* stream);

/* Device QLFC front end of one BWT block (host buffers): the coder-segment
 * split and, per segment, its symbol table and per-run symbol, rank and
 * length (arrays of n entries; tables 8 x 256). Returns the run count.
 * Counterpart: the front half of libbsc's QLFC coder (qlfc.cpp transform,
 * coder.cpp segment split), reached through bsc.rs:172-183. */
QZBSC_API int qzbsc_qlfc_front(const unsigned char* bwt, int n, int* nseg, int* seg_start, int* seg_size,
                               int* seg_runs, unsigned char* tables, int* nsym, unsigned char* run_sym,
                               unsigned char* run_rank, int* run_len);

/* Host-only QLFC coder stage of bsc_compress / bsc_decompress (the
 * "[nb][segment table][segments]" payload of one block from its BWT, and
 * back). No device needed. Counterpart: libbsc's coder.cpp
 * bsc_coder_compress / bsc_coder_decompress, reached through bsc.rs:172-183
 * and bsc.rs:421-429. */
QZBSC_API int qzbsc_coder_compress(const unsigned char* bwt, int n, unsigned char* out, int cap, int features);
QZBSC_API int qzbsc_coder_decompress(const unsigned char* in, int size, unsigned char* out, int n, int features);

/* Reads the size word and symbol table opening one coded segment. Test
 * hook with no reference counterpart: checks the segment framing against
 * libbsc-written blocks. */
QZBSC_API int qzbsc_segment_header(const unsigned char* seg, int size, int* n, unsigned char* table, int* nsym);

/* Serving loop with continuous batching (no counterpart in the reference;
 * it replaces qz's per-block rayon fan-out, mod.rs:402-404 and
 * decompress_impl.rs:117-120, for a long-running service).
 * create: HBM budget in bytes for the workspaces of batches in flight, a cap
 * on one batch's input bytes, and the worker count (one engine each). Every
 * worker reserves its workspaces for a full batch at creation; they never
 * grow after that.
 * submit: kind 0 = forward BWT (compress side: fills out, primary_out[b]
 * and, if aux_out != NULL, 256 ints of sampled indexes per block), kind 1 =
 * inverse BWT (decompress side: needs primary_in). The job is cut into work
 * units of whole blocks that workers batch with other jobs' units. Buffers
 * are host memory (pinned buffers from qzbsc_host_alloc copy fastest) and
 * must stay valid until wait returns. Returns a job id (> 0) or < 0
 * (LIBBSC_NOT_ENOUGH_MEMORY for a block larger than the batch cap or than
 * the 512 MiB chunk the forward sort workspace is sized for).
 * wait: blocks until every unit of the job is done; returns its status and,
 * optionally, its latency, its queueing time (until its first unit started)
 * and the most jobs any of its batches held. A job can be waited on once. */
QZBSC_API void* qzbsc_server_create(long long hbm_budget, long long max_batch_bytes, int workers);
/* As qzbsc_server_create (which uses a 32 MiB lane), with the small-job lane
 * set explicitly: one more worker takes only jobs of at most
 * small_lane_bytes input bytes, so small requests do not queue behind
 * batches of large ones (0 = no lane). Work units of 64 MiB, FIFO order. */
QZBSC_API void* qzbsc_server_create_lanes(long long hbm_budget, long long max_batch_bytes, int workers,
                                          long long small_lane_bytes);
/* As qzbsc_server_create_lanes, with the work-unit size (bytes of whole
 * blocks per unit; 0 = the batch cap) and the scheduling order: policy 0 =
 * oldest job first, 1 = the job with the fewest bytes not yet started first. */
QZBSC_API void* qzbsc_server_create_opts(long long hbm_budget, long long max_batch_bytes, int workers,
                                         long long small_lane_bytes, long long unit_bytes, int policy);
QZBSC_API void qzbsc_server_destroy(void* srv);
QZBSC_API int qzbsc_server_submit(void* srv, int kind, const unsigned char* in, unsigned char* out, const int* sizes,
                                  int nblocks, const int* primary_in, int* primary_out, int* aux_out);
QZBSC_API int qzbsc_server_wait(void* srv, int job, double* latency_ms, double* queue_ms, int* batch_jobs);
/* Largest footprint (bytes) the server had in flight. */
QZBSC_API long long qzbsc_server_peak_bytes(void* srv);
/* The batch cap the server settled on (input bytes per device batch). */
QZBSC_API long long qzbsc_server_batch_cap(void* srv);
/* Pinned (page-locked) host memory for request buffers, and its release.
 * No counterpart in the reference. */
QZBSC_API void* qzbsc_host_alloc(long long bytes);
QZBSC_API void qzbsc_host_free(void* p);
/* Free and total device memory (hipMemGetInfo), for admission checks and
 * tests of resident memory. No counterpart in the reference. */
QZBSC_API int qzbsc_device_memory(long long* free_bytes, long long* total_bytes);
/* Device bytes one batch of `bytes` input in `blocks` blocks needs:
 * kind 0 forward BWT, kind 1 inverse BWT. The forward's sort workspace
 * covers one chunk of whole blocks (512 MiB, QZBSC_CHUNK_MIB) and the
 * figure assumes no block is larger; only its staging grows with the
 * batch. */
QZBSC_API long long qzbsc_footprint(int kind, long long bytes, int blocks);

/* Radix digit-width tuner (no counterpart in the reference): which = 0 for
 * the round-0 suffix sort (candidates 8 and 11 bits), 1 for the sort of
 * groups larger than 4096 (8 and 9 bits). The first sorts of >= 4M
 * elements rotate through the candidates, timed with HIP events; then the
 * fastest per element is kept for the process (QZBSC_RADIX_BITS=8|9|11
 * forces one; the large-group sort takes at most 9). Writes the choice
 * (0 = undecided) and the mean ns per element of 8-, 9- and 11-bit digits
 * into ns_per_elem[3]; returns the number of trials. */
QZBSC_API int qzbsc_tuning(int which, int* choice, double* ns_per_elem);

/* Launch-shape tuner, every knob (no counterpart in the reference): knob 0
 * round-0 digit width (8, 11), 1 large-group digit width (8, 9), 2 tile
 * shape of the fused round-0 finish (QZBSC_FUSE_SHAPE codes: 0 = 512
 * threads per 4096-slot tile, 5 = 1024, both with the tile's suffixes in
 * LDS), 3 LDS budget of the small-group sort in bytes per workgroup (16384,
 * 0; it caps the workgroups sharing a CU). Each knob rotates its candidates
 * call by call on calls with enough work, timed with HIP events around its
 * own kernels, and keeps the fastest per unit of work (elements; for knob 3
 * small-group slots). Writes the choice (-1 while undecided), the number of
 * candidates, the candidates (cand[4]) and their mean ns per unit
 * (ns_per_unit[4]); returns the number of timed trials, or -1 for an
 * unknown knob. */
QZBSC_API int qzbsc_tune_knob(int knob, int* choice, int* ncand, int* cand, double* ns_per_unit);

/* HIP graph replay counts of this process (no reference counterpart).
 * Fixed launch sequences (the inverse BWT, round 0 of the forward sort) are
 * captured the second time in a row that an engine sees a batch shape and
 * replayed while the shape repeats; QZBSC_GRAPH=0 turns replay off. */
QZBSC_API int qzbsc_graph_stats(unsigned long long* captures, unsigned long long* replays);

/* LZP stage alone (host): the preprocessing bsc_compress applies when given
 * lzpHashSize / lzpMinLen (libbsc lzp.h bsc_lzp_compress / bsc_lzp_decompress;
 * parameter sets of crates/qz-lib/src/compression/bsc.rs:194-196, 237-246).
 * compress returns the encoded size, or LIBBSC_NOT_COMPRESSIBLE when it
 * would not be smaller than n (output holds n bytes); decompress writes
 * exactly outputSize bytes and returns 0 or LIBBSC_DATA_CORRUPT. */
QZBSC_API int qzbsc_lzp_compress(const unsigned char* input, unsigned char* output, int n, int lzpHashSize,
                                 int lzpMinLen);
QZBSC_API int qzbsc_lzp_decompress(const unsigned char* input, int inputSize, unsigned char* output, int outputSize,
                                   int lzpHashSize, int lzpMinLen);

/* quality_ctx coder (host): qz's context-adaptive quality compressor,
 * replacing compress_qualities_ctx / decompress_qualities_ctx
 * (crates/qz-lib/src/compression/quality_ctx.rs:296-391, 393-505) with the
 * same blob format and integer arithmetic. Reads are concatenated: read r
 * has qual_lens[r] quality bytes (Phred+33) and seq_lens[r] sequence bytes.
 * compress returns the blob size (0 for no reads), LIBBSC_BAD_PARAMETER for
 * the reference's input errors, or LIBBSC_NOT_ENOUGH_MEMORY if cap is too
 * small. decompress fills quals (room for the sum of seq_lens) and
 * qual_lens; returns 0 or LIBBSC_DATA_CORRUPT. */
QZBSC_API long long qzbsc_quality_ctx_compress(const unsigned char* quals, const int* qual_lens,
                                               const unsigned char* seqs, const int* seq_lens, int num_reads,
                                               unsigned char* out, long long cap);
QZBSC_API int qzbsc_quality_ctx_decompress(const unsigned char* data, long long size, const unsigned char* seqs,
                                           const int* seq_lens, int num_reads, unsigned char* quals, int* qual_lens);

/* Pool control: at most n engines (streams + resident workspaces). */
QZBSC_API int qzbsc_set_max_engines(int n);
/* Releases the resident workspaces of every idle engine and the cached
 * pinned host buffers of the front end's downloads (no reference
 * counterpart: for long-running processes between workloads). Returns the
 * bytes released. */
QZBSC_API long long qzbsc_trim(void);
/* 1 if a HIP device is usable, else 0. */
QZBSC_API int qzbsc_gpu_available(void);

/* Memory-system calibration (no counterpart in the reference; not on the
 * block path): the mean ms per launch of `iters` launches of one access
 * shape over a device array of array_bytes bytes (< 16 GiB). mode 0: items
 * random 4-byte gathers dst[i] = src[idx[i]] with idx and dst streamed;
 * mode 1: items random 4-byte stores dst[idx[i]] = i; mode 2: a 16-byte
 * vector copy of the whole array (items unused). Indexes are uniform over
 * the array. Allocates and frees its own buffers. */
QZBSC_API int qzbsc_calib_access(int mode, long long array_bytes, long long items, int iters, double* ms);

#ifdef __cplusplus
}
#endif
#endif
