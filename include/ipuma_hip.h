/*
 * ipuma_hip.h -- C ABI of the MI355X batched Smith-Waterman (affine gap) engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes, no
 * exceptions across the boundary.  Every entry point returns IPUMA_OK (0) or
 * a negative IPUMA_E* code; ipuma_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * Each entry point cites the reference interface it replaces
 * (xiamaz/ipuma-lib, file:line).  Results use the reference's packing
 * (src/codelets/codelet_batch_affine.cpp:117-123):
 *     a_range[k] = (uint16 Aend << 16) | uint16 Astart
 *     b_range[k] = (uint16 Bend << 16) | uint16 Bstart
 * With forward_only the starts are 0, as in the reference.
 */
#ifndef IPUMA_HIP_H
#define IPUMA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPUMA_ABI_VERSION 1

enum {
  IPUMA_OK = 0,
  IPUMA_EINVAL = -1,    /* bad argument (size, pointer, enum, alignment) */
  IPUMA_ENOMEM = -2,    /* device or host allocation failed / over budget */
  IPUMA_EDEVICE = -3,   /* HIP runtime or kernel failure */
  IPUMA_ECONFIG = -4,   /* configuration the reference rejects (e.g. `simple`) */
  IPUMA_ETOOLONG = -5,  /* sequence longer than the engine supports */
  IPUMA_EBUSY = -6,     /* admission refused (request server) */
  IPUMA_ECLOSED = -7    /* request server shut down */
};

/* swatlib::Similarity (src/swatlib/similarity.h:108-113). */
enum { IPUMA_SIM_NA = 0, IPUMA_SIM_BLOSUM50 = 1, IPUMA_SIM_BLOSUM62 = 2, IPUMA_SIM_SIMPLE = 3 };
/* swatlib::DataType (src/swatlib/encoding.h:12-16). */
enum { IPUMA_DT_NA = 0, IPUMA_DT_AA = 1, IPUMA_DT_STRING = 2 };
/* ipu::VertexType (src/types.h:19) and ipu::Algorithm (src/types.h:21). */
enum { IPUMA_VT_CPP = 0, IPUMA_VT_ASSEMBLY = 1, IPUMA_VT_MULTI = 2, IPUMA_VT_MULTIASM = 3 };
enum { IPUMA_FILL_FIRST = 0, IPUMA_ROUND_ROBIN = 1, IPUMA_GREEDY = 2 };

/* Mirrors ipu::SWConfig field for field (src/ipu_config.h:11-19). */
typedef struct ipuma_sw_config {
  int32_t gap_init;        /* gapInit   (default 0)  */
  int32_t gap_extend;      /* gapExtend (default -1) */
  int32_t match_value;     /* matchValue (1)         */
  int32_t mismatch_value;  /* mismatchValue (-1)     */
  int32_t ambiguity_value; /* ambiguityValue (-1)    */
  int32_t similarity;      /* IPUMA_SIM_*            */
  int32_t datatype;        /* IPUMA_DT_*             */
} ipuma_sw_config;

/* Mirrors ipu::IPUAlgoConfig (src/ipu_config.h:21-31).  On MI355X the tile
 * fields are capacity hints only: tiles_used*max_batches bounds the number of
 * comparisons of one compare_local call (ipu_batch_affine.cpp:410-416) and
 * max_ab bounds |A| (:418-423).  vtype / fill_algo / use_remote_buffer /
 * transmission_programs / io_tiles are accepted and ignored. */
typedef struct ipuma_algo_config {
  int32_t tiles_used;
  int32_t max_ab;
  int32_t max_batches;
  int32_t bufsize;
  int32_t vtype;
  int32_t fill_algo;
  int32_t forward_only;
  int32_t use_remote_buffer;
  int32_t transmission_programs;
  int32_t io_tiles;
} ipuma_algo_config;

typedef struct ipuma_engine ipuma_engine;

/* Fill the reference defaults (src/ipu_config.h:11-31). */
void ipuma_default_sw_config(ipuma_sw_config *c);
void ipuma_default_algo_config(ipuma_algo_config *c);

/* Replaces SWAlgorithm(SWConfig, IPUAlgoConfig[, thread_id, slotCap, ipuCount,
 * runExecutor]) (src/ipu_batch_affine.h:76-77, .cpp:371-403): binds the
 * current HIP device, builds the scoring table, creates the engine stream.
 * `simple` similarity fails with IPUMA_ECONFIG like selectMatrix does
 * (src/swatlib/similarity.cpp:100-101). */
int ipuma_engine_create(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                        ipuma_engine **out);
/* Replaces ~SWAlgorithm (src/ipu_batch_affine.cpp:1014-1022). */
void ipuma_engine_destroy(ipuma_engine *e);

/* Last error message of the calling thread ("" if none). */
const char *ipuma_last_error(void);

/*
 * Replaces compare_local(A, B) + get_result()
 * (src/ipu_batch_affine.h:96,99; .cpp:564-663, :408).
 * Comparison k aligns A_k = a_chars[a_offsets[k] .. a_offsets[k+1]) against
 * B_k = b_chars[b_offsets[k] .. b_offsets[k+1]).  Host buffers in and out;
 * results are written for k < n.  Errors like the reference: more than
 * tiles_used*max_batches comparisons or |A_k| > max_ab -> IPUMA_EINVAL /
 * IPUMA_ETOOLONG (the reference throws / exits).
 */
int ipuma_compare_local(ipuma_engine *e, int32_t n,
                        const char *a_chars, const int64_t *a_offsets,
                        const char *b_chars, const int64_t *b_offsets,
                        int32_t *scores, int32_t *a_range, int32_t *b_range);

/*
 * Replaces compare_mn_local(Seqs, Cmps) + get_result()
 * (src/ipu_batch_affine.h:100; .cpp:697-750): comparison k aligns
 * Seqs[cmps[2k]] (A) against Seqs[cmps[2k+1]] (B); ipu::Comparison is the
 * packed {int32 indexA, int32 indexB} pair (src/types.h:12-15).  Each distinct
 * sequence is encoded and uploaded once.
 */
int ipuma_compare_mn_local(ipuma_engine *e, int32_t nseqs, const char *chars,
                           const int64_t *offsets, int32_t ncmps, const int32_t *cmps,
                           int32_t *scores, int32_t *a_range, int32_t *b_range);

/*
 * Device-resident batch: the hot path with inputs already in HBM (what the
 * device loop of src/ipu_batch_affine.cpp:340-365 consumes).  `codes` holds
 * encoded residues (ipuma_encode_device or the host encoder); pair k aligns
 * codes[a_off[k] .. +a_len[k]) against codes[b_off[k] .. +b_len[k]).  Offsets
 * must be multiples of 4, every sequence must lie inside [0, codes_len) and
 * `codes` must be readable up to the next multiple of 4 past every sequence.
 * All pointers are device pointers; the call enqueues on `stream` (a
 * hipStream_t, NULL = the engine's stream) and returns without waiting;
 * ipuma_synchronize reports a length error (> 65535, positions are uint16).
 * Codes outside the matrix are not allowed (ipuma_encode_device never makes
 * them).
 */
int ipuma_compare_device(ipuma_engine *e, int32_t n, const uint8_t *codes, int64_t codes_len,
                         const int64_t *a_off, const int32_t *a_len,
                         const int64_t *b_off, const int32_t *b_len,
                         int32_t *scores, int32_t *a_range, int32_t *b_range,
                         void *stream);

/*
 * ipuma_compare_device that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments and results as for ipuma_compare_device, forward pass only (the
 * engine must be forward-only, else IPUMA_EINVAL).  Its device work space is a
 * static array of about 260 MiB per GPU (held by every process that loads the
 * library, until it exits), which bounds one call to n <= 65536 comparisons
 * (IPUMA_EINVAL beyond) and |A| <= 4096 (a longer A is reported as
 * IPUMA_ETOOLONG by ipuma_synchronize, and its score is written as -1).
 * Calls on one GPU share that array: calls issued outside graph capture are
 * ordered on the device in issue order; graph replays, and calls captured
 * into graphs, that overlap on one GPU race.  The packed kernel runs on at
 * most 256 blocks, the int32 kernel on 256 waves.
 */
int ipuma_compare_device_noalloc(ipuma_engine *e, int32_t n, const uint8_t *codes,
                                 int64_t codes_len, const int64_t *a_off, const int32_t *a_len,
                                 const int64_t *b_off, const int32_t *b_len, int32_t *scores,
                                 int32_t *a_range, int32_t *b_range, void *stream);

/*
 * One-vs-many search (the database-search use of compare_local, with every
 * query against every database sequence): comparison q * ndb + d aligns
 * database sequence d (A) against query q (B); all sequences are in one
 * device arena of codes (offsets 4-aligned, as for ipuma_compare_device).
 * Writes every score (scores: nq * ndb, or NULL) and per query the top k
 * (k <= 1024) database indices and scores, ordered by score descending, then
 * index ascending (top_idx / top_score: nq * k; -1 / 0 past ndb).  Forward
 * pass only.  No reference counterpart: the reference expresses search as
 * compare_local / compare_mn_local over explicit pairs (src/ipu_batch_affine.h:99-100).
 */
int ipuma_search_device(ipuma_engine *e, const uint8_t *codes, int64_t codes_len, int32_t nq,
                        const int64_t *q_off, const int32_t *q_len, int32_t ndb,
                        const int64_t *db_off, const int32_t *db_len, int32_t k,
                        int32_t *scores, int32_t *top_idx, int32_t *top_score, void *stream);

/*
 * Prepared input buffers: the reference's wire format, byte for byte
 * (src/ipu_batch_affine.cpp:752-950, src/ipu_config.cpp:6-14).  An input
 * buffer holds tiles_used tile regions of one-byte residue codes, then per
 * tile max_batches x {|A|, A offset, |B|, B offset} int32 metadata (offsets
 * inside the tile's region); comparison k lands in slot mapping[k] = tile *
 * max_batches + position.  The results buffer is [scores | a ranges |
 * b ranges], one int32 per slot.
 */
/* IPUAlgoConfig::getInputBufferSize32b (src/ipu_config.cpp:14): words of an
 * input buffer for this config (negative on error). */
int64_t ipuma_input_buffer_words(const ipuma_algo_config *algo);
/* Replaces SWAlgorithm::prepare_remote (src/ipu_batch_affine.h:114,
 * .cpp:891-950): the algo config's fill algorithm places comparisons k = A_k
 * vs B_k (strings as in ipuma_compare_local), the buffer is cleared and
 * written, mapping[k] receives the slot of comparison k and *max_bucket (may
 * be NULL) the largest tile fill in bytes.  Host only. */
int ipuma_prepare_remote(const ipuma_sw_config *sw, const ipuma_algo_config *algo, int32_t n,
                         const char *a_chars, const int64_t *a_offsets, const char *b_chars,
                         const int64_t *b_offsets, int32_t *inputs, int64_t input_words,
                         int32_t *mapping, int32_t *max_bucket);
/* The preparation step of compare_mn_local (fillMNBuckets + fill_input_buffer,
 * src/ipu_batch_affine.cpp:697-705, 752-822): comparisons cmps[2k] vs
 * cmps[2k+1] over one sequence list, a sequence stored once per tile.  The
 * buffer is cleared first.  Host only. */
int ipuma_prepare_remote_mn(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                            int32_t nseqs, const char *chars, const int64_t *offsets,
                            int32_t ncmps, const int32_t *cmps, int32_t *inputs,
                            int64_t input_words, int32_t *mapping, int32_t *max_bucket);
/* Replaces SWAlgorithm::transferResults (src/ipu_batch_affine.h:115-116,
 * .cpp:674-695): results of comparison k from slot mapping[k] of each third
 * of the results buffer (result_words must be a multiple of 3).  Host only. */
int ipuma_transfer_results(const int32_t *results, int64_t result_words, const int32_t *mapping,
                           int32_t n, int32_t *scores, int32_t *a_range, int32_t *b_range);
/* Replaces prepared_remote_compare (src/ipu_batch_affine.h:105,
 * .cpp:557-562): runs every slot of a prepared buffer (laid out for this
 * engine's algo config; `lot` is the layout lot of transmission_programs > 1,
 * else 0) and writes the first result_words words of [scores | a ranges |
 * b ranges].  Empty slots score 0.  Blocks until the results are written. */
int ipuma_prepared_compare(ipuma_engine *e, const int32_t *inputs, int64_t input_words,
                           int32_t *results, int64_t result_words, int32_t lot);
/* Replaces async_submit_prepared_remote_compare /
 * blocking_join_prepared_remote_compare (src/ipu_batch_affine.h:107-108,
 * .cpp:496-555): the engine's executor thread runs submitted buffers in
 * order (layout lot 0); both buffers must stay valid until ipuma_prepared_join
 * returns, which reports the job's status.  Any thread may submit or join. */
int ipuma_prepared_submit(ipuma_engine *e, const int32_t *inputs, int64_t input_words,
                          int32_t *results, int64_t result_words, int64_t *job);
int ipuma_prepared_join(ipuma_engine *e, int64_t job);
/* Times of a prepared job: the input buffer's copy to HBM and the residue
 * gather plus device batch before the results copy (HIP events on the
 * engine's stream), and the executor's wall time for the whole job (copies
 * in and out included, queueing excluded). */
typedef struct ipuma_job_times {
  double h2d_ms;
  double device_ms;
  double run_ms;
} ipuma_job_times;
/* ipuma_prepared_join that also reports the job's times (`times` may be
 * NULL).  Replaces what a joined Job carries: Job::h2dCycles / innerCycles
 * (src/ipu_batch_affine.h:41-48, read from the device program's cycle
 * counters at src/ipu_batch_affine.cpp:123-127, 345-351) and
 * SubmittedBatch::runTick (:28-38, ticked at :81 and :127); multiplied by
 * ipuma_device_clock_hz the times give cycles of the same meaning. */
int ipuma_prepared_join_timed(ipuma_engine *e, int64_t job, ipuma_job_times *times);
/* The device's peak shader clock in Hz (hipDeviceProp_t::clockRate).
 * Replaces IPUAlgorithm::getTileClockFrequency (src/ipu_base.h:50,
 * src/ipu_base.cpp:197-199), which converts the reference's cycle counts to
 * seconds. */
int ipuma_device_clock_hz(ipuma_engine *e, double *hz);

/*
 * Traceback (no counterpart in the reference's IPU path; its CPU comparator,
 * SSW, reports one: extern/include/ssw/ssw.cpp:57-91, ssw_core.cpp:854-1053).
 * For comparison k of a compare with starts (scores and ranges as returned),
 * the CIGAR from the start to the end of its alignment: runs (length << 4 |
 * op), op 0 = M (A and B), 1 = I (A only), 2 = D (B only), written to
 * cigar[cigar_offsets[k] ..], their count to cigar_len[k] (0 for a score of
 * 0).  Capacity cigar_offsets[k+1] - cigar_offsets[k] must be at least
 * (Aend - Astart + 1) + (Bend - Bstart + 1).  A[Astart..Aend] and
 * B[Bstart..Bend] are aligned globally (their optimum is the local score);
 * ties prefer the diagonal, then the A gap when strictly better, and a gap
 * opens only when strictly better than extending.  Host buffers in and out.
 */
int ipuma_traceback_local(ipuma_engine *e, int32_t n, const char *a_chars,
                          const int64_t *a_offsets, const char *b_chars, const int64_t *b_offsets,
                          const int32_t *scores, const int32_t *a_range, const int32_t *b_range,
                          uint32_t *cigar, const int64_t *cigar_offsets, int32_t *cigar_len);
/* Of the last ipuma_traceback_local call: its region cells (sum of
 * (Aend - Astart + 1) * (Bend - Bstart + 1) over comparisons with a positive
 * score) and, with timing on (ipuma_set_timing), the time of its traceback
 * kernels (HIP events around each launch on the engine's stream, summed), else
 * 0.  Either pointer may be NULL.  No reference counterpart: the traceback's
 * throughput. */
int ipuma_last_traceback_stats(ipuma_engine *e, int64_t *region_cells, double *kernel_ms);

/* Batching policy of the device path (replaces IPUAlgoConfig.fillAlgo's
 * partition::fillFirst/roundRobin/greedy, src/partition.cpp:197-462). */
enum {
  IPUMA_BATCH_BUCKETED = 0,  /* device length bucketing + persistent queue (default) */
  IPUMA_BATCH_PLAIN = 1,     /* one pair per lane, input order, padded to the batch max */
  IPUMA_BATCH_PLAIN_TILE = 2 /* one pair per lane, input order, each wave's group padded
                                to its own longest pair */
};
/* Precision policy of the device path. */
enum {
  IPUMA_PREC_AUTO = 0,  /* packed int16, exact int32 re-run of saturated pairs */
  IPUMA_PREC_INT32 = 1  /* int32 only */
};
int ipuma_set_batching(ipuma_engine *e, int32_t mode);
int ipuma_set_precision(ipuma_engine *e, int32_t mode);
int ipuma_set_forward_only(ipuma_engine *e, int32_t forward_only);
/* Lanes per comparison of the packed kernel: 0 = chosen per comparison by the
 * bucketing (default), or one of 1, 2, 4, 8, 16, 64 for every comparison.
 * No reference counterpart (the IPU runs one comparison per worker). */
int ipuma_set_lanes_per_pair(ipuma_engine *e, int32_t lanes);
/* Groups per resident wave the length bucketing aims for (1..64, default 1):
 * more rounds give the longest-first schedule finer pieces to balance, at the
 * price of deeper pipelines (more row padding) per comparison.  No reference
 * counterpart (the IPU's work split is fixed by its tile mapping). */
int ipuma_set_work_rounds(ipuma_engine *e, int32_t rounds);
/* Scale of the bucketing's cap on a group's steps (0.25..4, default 1): above
 * 1, fewer and longer groups (shallower pipelines, less row padding); below 1,
 * more and shorter groups (a finer longest-first queue).  The knob the
 * autotuner (ipuma_amd.autotune) searches together with the rounds.  No
 * reference counterpart (the IPU's work split is fixed by its tile mapping). */
int ipuma_set_step_cap_scale(ipuma_engine *e, float scale);

/* Encode characters to residue codes on the device with the engine's
 * encoding (swatlib Encoding, src/swatlib/encoding.cpp:40-62). */
int ipuma_encode_device(ipuma_engine *e, const char *d_chars, int64_t len,
                        uint8_t *d_codes, void *stream);

/* Statistics of the last device batch (for logging / the bench). */
typedef struct ipuma_batch_stats {
  int64_t pairs;           /* comparisons */
  int64_t cells;           /* sum |A|*|B| */
  int64_t padded_cells;    /* cells the kernels computed, padding included */
  int64_t groups;          /* work groups of the persistent kernel */
  int64_t rerun_pairs;     /* pairs re-run in int32 after int16 saturation */
  int32_t lanes_per_pair_max;
  int32_t rows_per_lane;
  int32_t band_groups;     /* whole-wave groups shared by the waves of a block */
  int32_t head_groups;     /* longest other groups, started with the band groups */
  double fast_kernel_ms;   /* packed kernel, HIP events on its stream (timing on) */
  double batch_ms;         /* whole device batch, same events (timing on) */
} ipuma_batch_stats;
/* Record HIP events around the packed kernel and the whole device batch of the
 * following calls (read back by ipuma_last_batch_stats). */
int ipuma_set_timing(ipuma_engine *e, int32_t on);
int ipuma_last_batch_stats(ipuma_engine *e, ipuma_batch_stats *out);

/* Measured throughput of the VALU instructions the DP issues, on the current
 * device: packed int16 saturating add + max (v_pk_add_i16 / v_pk_max_i16, 2
 * ops per lane-instruction) and 32-bit add + max, in tera-ops per second
 * (either pointer may be NULL).  Runs two short kernels (about 0.3 s).  No
 * reference counterpart: the calibrated peak of the packed kernel's
 * roofline, reported beside the nominal 78.6 T (256 CUs x 4 SIMDs x 16 lanes
 * x 2.4 GHz x 2). */
int ipuma_valu_peak(double *pk_int16_tops, double *int32_tops);

/* Synchronize the engine stream. */
int ipuma_synchronize(ipuma_engine *e);

/*
 * Request server: replaces the reference's batch loop over asynchronous jobs
 * (IPUMultiDriver, ipusw/ipumultidriver.hpp:10-145; run_comparison,
 * ipusw/run_comparison.hpp:67-172; async_submit_prepared_remote_compare /
 * blocking_join_prepared_remote_compare, src/ipu_batch_affine.h:105-108).
 * Requests of any size from any host thread are encoded and copied once into
 * an HBM arena (admitted against hbm_budget_bytes), merged with other pending
 * requests into one device batch, and retired in order; wait() returns a
 * request's results (packed like compare_local's) and its latency from
 * submit to results on the host.
 */
typedef struct ipuma_server ipuma_server;
typedef struct ipuma_server_config {
  int64_t hbm_budget_bytes; /* arena for resident request sequences (default 64 GiB) */
  int32_t max_batch_pairs;  /* most comparisons per device batch, and per request (65536) */
  int32_t max_wait_us;      /* a pending request waits at most this long while the GPU is
                               busy before its batch is launched (2000) */
  int32_t max_inflight;     /* device batches enqueued at once, 1..8 (2) */
  int32_t block_on_full;    /* 1: submit waits for arena space; 0: it returns IPUMA_EBUSY */
  int32_t merge_requests;   /* 1: pending requests share device batches (default); 0: one
                               device batch per request (the unbatched baseline) */
} ipuma_server_config;
typedef struct ipuma_server_stats {
  int64_t requests;             /* admitted */
  int64_t batches;              /* retired device batches */
  int64_t pairs;                /* comparisons retired */
  int64_t cells;                /* sum |A|*|B| retired */
  int64_t max_batch_pairs_seen; /* largest merged batch */
  double busy_ms;               /* sum of batch launch-to-retire times */
  int64_t arena_used_bytes;     /* HBM arena held by requests not yet retired */
  int64_t pending_pairs;        /* comparisons waiting for a batch */
  double span_ms;               /* first batch launch to last retire */
} ipuma_server_stats;
void ipuma_default_server_config(ipuma_server_config *c);
/* One engine (sw, algo: scoring and forward_only) plus its dispatcher and
 * completer threads on the current HIP device. */
int ipuma_server_create(const ipuma_sw_config *sw, const ipuma_algo_config *algo,
                        const ipuma_server_config *cfg, ipuma_server **out);
/* Retires everything submitted, joins the threads, frees the arena.  No call
 * on the server may be in progress. */
void ipuma_server_destroy(ipuma_server *s);
/* Comparison k aligns A_k = a_chars[a_offsets[k] .. a_offsets[k+1]) against
 * B_k likewise; the caller's buffers may be reused when this returns.  Fills
 * *ticket (also for an upload that failed, whose wait reports the error). */
int ipuma_server_submit(ipuma_server *s, int32_t n, const char *a_chars,
                        const int64_t *a_offsets, const char *b_chars,
                        const int64_t *b_offsets, int64_t *ticket);
/* Blocks until the request is retired; copies its n results (any output may
 * be NULL) and its latency; the ticket is then released. */
int ipuma_server_wait(ipuma_server *s, int64_t ticket, int32_t *scores, int32_t *a_range,
                      int32_t *b_range, double *latency_ms);
int ipuma_server_get_stats(ipuma_server *s, ipuma_server_stats *out);
/* The server's engine (for its batching / precision knobs; not for batches). */
int ipuma_server_engine(ipuma_server *s, ipuma_engine **out);

#ifdef __cplusplus
}
#endif
#endif /* IPUMA_HIP_H */
