/* extremz.h -- C ABI of the MI355X ExtreM correction engine.
 *
 * Plain pointers and sizes only.  Fields are raw little-endian arrays with x
 * fastest (idx = x + W*(y + H*z)), the layout of decp.bin / fixed.bin
 * (src/ExtreM_cuda.cu:1002, src/ExtreM.cpp:2918-2928).  All arithmetic is
 * IEEE float64 without contraction, as the reference (src/ExtreM.cpp:33).
 *
 * Every entry point returns 0 or a negative EXTREMZ_E* code.  A context owns
 * one HIP stream and the device work buffers; calls on one context must not
 * overlap, calls on different contexts may.
 */
#ifndef EXTREMZ_H
#define EXTREMZ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXTREMZ_OK 0
#define EXTREMZ_ENOTCONV (-3)  /* an iteration cap was reached */
#define EXTREMZ_EHIP (-5)      /* a HIP runtime call failed */
#define EXTREMZ_EAGAIN (-11)   /* a served job is not finished yet */
#define EXTREMZ_ENOMEM (-12)   /* device allocation failed */
#define EXTREMZ_ENODEV (-19)   /* no usable gfx950 device */
#define EXTREMZ_EINVAL (-22)   /* bad shape, bound, dtype or non-finite input */
#define EXTREMZ_ENOSYS (-38)   /* libzstd.so.1 is not available */

#define EXTREMZ_DTYPE_F32 0
#define EXTREMZ_DTYPE_F64 1

/* Indices into extremz_stats.counts. */
enum {
  EXTREMZ_STAT_STATUS = 0,
  EXTREMZ_STAT_C_LOOP_ITERS,    /* iterations of the c_loops inner while */
  EXTREMZ_STAT_OUTER_ITERS,     /* iterations of main.cpp:277 while */
  EXTREMZ_STAT_R_LOOP_CALLS,
  EXTREMZ_STAT_S_LOOP_CALLS,
  EXTREMZ_STAT_SADDLE_LOOP_CALLS,
  EXTREMZ_STAT_C_LOOPS_CALLS,
  EXTREMZ_STAT_EDITED,          /* vertices with a non-zero edit count */
  EXTREMZ_STAT_N_MAX,
  EXTREMZ_STAT_N_MIN,
  EXTREMZ_STAT_N_SADDLE1,
  EXTREMZ_STAT_N_SADDLE2,
  EXTREMZ_STAT_LABEL_PASSES,    /* full-grid manifold label computations */
  EXTREMZ_STAT_INC_PASSES,      /* CP passes of decp done incrementally (no full-grid pass) */
  EXTREMZ_STAT_HOST_SYNCS,      /* device->host counter reads */
  EXTREMZ_STAT_FRONTIER_ITEMS,  /* false-critical-point frontier entries */
  EXTREMZ_STAT_COUNT_ = 16
};

/* Kernel families timed with HIP events on the context stream.  Each event
 * pair brackets one launch of one kernel, except the families marked
 * "sequence", whose event pairs bracket the short kernel sequences named. */
enum {
  EXTREMZ_K_CLASSIFY = 0, /* k_classify: full-grid classification (extremz_classify) */
  EXTREMZ_K_CHECK_FULL,   /* k_mask_cols: full-grid false-critical-point rescan */
  EXTREMZ_K_CHECK_INCR,   /* k_check_list / k_check_two: frontier and edited-neighbourhood
                             re-checks */
  EXTREMZ_K_LABEL_TILE,   /* k_labels_tile: steepest pointers + lower masks, resolved
                             inside LDS tiles */
  EXTREMZ_K_INCREMENTAL,  /* sequence: incremental CP pass of decp (pointer, root and
                             saddle updates) */
  EXTREMZ_K_LABEL_CHASE,  /* k_chase: full-grid walk of the remaining pointers to roots */
  EXTREMZ_K_SADDLE_PASS,  /* k_saddle_list (decp mode): triplet ends, candidates,
                             changed saddles */
  EXTREMZ_K_SLOT_GATHER,  /* sequence: largest / smallest saddle per extremum (gather,
                             merge, ties) */
  EXTREMZ_K_PROPOSE,      /* sequence: correction proposals (all loops) */
  EXTREMZ_K_APPLY,        /* k_apply: apply_delta over touched vertices */
  EXTREMZ_K_RESERVE,      /* sequence: wrong-neighbour reservation rounds */
  EXTREMZ_K_SCAN_PROPOSE, /* k_scan_propose: the first full false-critical-point scan of
                             a call, which makes the frontier's proposals */
  EXTREMZ_K_SADDLE_ORIG,  /* k_saddle_list (original mode): triplet ends and the
                             extremum-saddle adjacency counts of the original field */
  EXTREMZ_K_CSR_FILL,     /* k_csr_fill: the extremum -> saddle index of the original */
  EXTREMZ_K_FAMILIES_ = 16
};

typedef struct extremz_opts {
  int max_c_iters; /* cap on c_loops inner iterations per run; 0 = 100000 */
  int max_outer;   /* cap on outer iterations; 0 = 10000 */
  int timing;      /* bit f set: time family f with HIP events (0 = none) */
  int fixed_step_q; /* 0: the CPU path's midpoint / snap rules (src/ExtreM.cpp);
                       q in [1, 254]: the GPU fork's fixed-step rule set
                       (edit_type 0, src/ExtreM_cuda.cu:2905, 3065-3076): every
                       proposal is -xi/q, an edit count reaching q - 1 snaps
                       to orig - xi with count q */
  int adjacency;    /* EXTREMZ_ADJ_STENCIL14 (0): the 14-neighbour stencil of
                       src/ExtreM_cuda.cu:155-162 for every field;
                       EXTREMZ_ADJ_PLANE4 (1): the CPU path's plane adjacency,
                       maxNeighbors = 6 (main.cpp:20) over the direction table
                       of src/ExtreM.cpp:16-21, i.e. +-x and +-y, with the
                       triangle links of src/ExtreM.cpp:597-627, 703-891;
                       fields of depth 1 only (else EXTREMZ_EINVAL) */
  int rule_set;     /* EXTREMZ_RULES_CPU (0): the rules above;
                       EXTREMZ_RULES_F32 (1): the float32 fork's rule set
                       (src/ExtreM_cuda_f.cu:2583-2600, 2703, 2749, 2804,
                       2854-2866) within the same control flow: xi, orig and
                       decp in float32 (inputs float32 cannot hold exactly are
                       EXTREMZ_EINVAL), every fix proposal -xi/4 and the star
                       fixes (orig - xi) - decp in float32, an edit applied
                       while |delta| > 1e-6, count < 6 and
                       |orig - (decp + delta)| < xi, else decp := orig - xi
                       with count 7.  Needs fixed_step_q 0. */
  double persistence; /* > 0: the GPU fork's persistence threshold thr
                       (findAscPaths, src/ExtreM_cuda.cu:1877-1881; thr =
                       (max - min) * sim, :231): a 2-saddle's triplet keeps
                       only the maxima with |f(max) - f(saddle)| >= thr, in
                       each field's own values.  Absolute, like xi; 0 = off. */
} extremz_opts;

#define EXTREMZ_ADJ_STENCIL14 0
#define EXTREMZ_ADJ_PLANE4 1
#define EXTREMZ_RULES_CPU 0
#define EXTREMZ_RULES_F32 1

typedef struct extremz_stats {
  int64_t counts[EXTREMZ_STAT_COUNT_];
  double kernel_ms[EXTREMZ_K_FAMILIES_];     /* summed event time */
  int64_t kernel_launches[EXTREMZ_K_FAMILIES_];
  double wall_ms;                            /* host wall time of the call */
} extremz_stats;

typedef struct extremz_ctx extremz_ctx;

/* Creates a context on a HIP device.  No reference counterpart: the
 * reference keeps all state in globals (include/api/ExtreM.h:17-41). */
int extremz_ctx_create(int device, extremz_ctx** out);
int extremz_ctx_destroy(extremz_ctx* ctx);

/* The correction call: replaces the sequence of main.cpp:65-446 (getdata's
 * field + bound, pre-snap, classifyAllVertices, manifolds, computeCP,
 * c_loops / r_loops / s_loops / saddle_loops until every counter is zero)
 * whose result main.cpp:507 writes to fixed.bin.  orig is float32 or
 * float64 (orig_dtype), promoted exactly; decp is edited in place; xi is
 * the absolute bound (src/ExtreM.cpp:138: (max-min)*er).  counts (nullable)
 * receives delta_counter (main.cpp:57) as uint8.  Host pointers. */
int extremz_correct(const void* orig, int orig_dtype, double* decp, int W, int H,
                    int D, double xi, const extremz_opts* opts,
                    extremz_stats* stats, uint8_t* counts);

/* Same on device-resident buffers (orig float64), on the context stream.
 * Returns when the result is in d_decp. */
int extremz_correct_device(extremz_ctx* ctx, const double* d_orig, double* d_decp,
                           int W, int H, int D, double xi,
                           const extremz_opts* opts, extremz_stats* stats,
                           uint8_t* d_counts);

/* Phase (a): vertex types (0 min, 1 1-saddle, 2 2-saddle, 3 both, 4 max,
 * 5 regular; src/ExtreM.cpp:880) and lower-star direction masks of a float64
 * field; replaces classifyAllVertices (src/ExtreM.cpp:1088-1097) with the
 * LUT rule of src/ExtreM_cuda.cu:3796-3853.  Host pointers. */
int extremz_classify(const double* field, int W, int H, int D, uint8_t* types,
                     uint16_t* lower_masks);

/* Phase (b): fully compressed steepest-ascent (ds) and steepest-descent (as)
 * root labels; replaces ComputeDescendingManifold / ComputeAscendingManifold
 * (src/ExtreM.cpp:144-183, 502-541).  Host pointers. */
int extremz_labels(const double* field, int W, int H, int D, int32_t* ds,
                   int32_t* as);

/* Phases (a) and (b) of the original field in one tile pass, as the
 * correction computes them (main.cpp:85-89: classifyAllVertices, then
 * ComputeDescendingManifold / ComputeAscendingManifold): the labels of
 * extremz_labels with the types and lower masks of extremz_classify, the
 * latter from the tile pass's own stencil scan.  Host pointers. */
int extremz_label_pass(const double* field, int W, int H, int D, int32_t* ds, int32_t* as,
                       uint8_t* types, uint16_t* lower_masks);

/* The classification table in the layout of LUT.bin (24 800 int32 over the
 * 27 boundary keys of src/ExtreM_cuda.cu:84-110), generated from the link
 * definition; loadLUTToGPU (src/ExtreM_cuda.cu:244-260) reads the same data
 * from disk.  Returns the entry count (or the required size if cap is too
 * small).  Runs on the host; no GPU needed. */
int extremz_lut(int32_t* out, int cap);

/* Phase (d): edit encoding, cost() of src/ExtreM.cpp:272-389.
 * index_bitmap receives the 3-bit LSB-first packing of per-vertex edit
 * counts (0 for unedited vertices; encodeTo3BitBitmap, src/ExtreM.cpp:220-250),
 * ceil(3N/8) bytes.  For every vertex whose count is 6 (snapped) the edit
 * -(decp0 - (orig - xi)) is written in index order to edits (float64), its
 * 52-bit mantissa to mantissas and unbiased exponent to exponents (the split
 * of src/ExtreM_cuda.cu:4466-4476).  n_edits receives the number of edits;
 * the three edit arrays (nullable) need room for N entries.  Host pointers;
 * runs on the GPU. */
int extremz_encode(const double* orig, const double* decp_before,
                   const double* decp_after, const uint8_t* counts, int64_t n,
                   double xi, uint8_t* index_bitmap, double* edits,
                   uint64_t* mantissas, int32_t* exponents, int64_t* n_edits);

/* No counterpart in the reference: extremz_encode for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same outputs as extremz_encode, on
 * device pointers, queued on ctx's stream; it returns without waiting (read
 * *d_n_edits, an int64 in device memory, after synchronising the stream).
 * It allocates nothing: its only scratch is a static device array of
 * per-block edit counts (4 MiB on each device, held until the process ends)
 * shared by every call on a device, so n is at most
 * EXTREMZ_ENCODE_NOALLOC_MAX_N (else EXTREMZ_EINVAL) and calls that overlap
 * on one device race.  The edit arrays (nullable) need room for n entries. */
#define EXTREMZ_ENCODE_NOALLOC_MAX_N ((int64_t)1 << 30)
int extremz_encode_noalloc(extremz_ctx* ctx, const double* d_orig, const double* d_decp_before,
                           const double* d_decp_after, const uint8_t* d_counts, int64_t n,
                           double xi, uint8_t* d_index_bitmap, double* d_edits,
                           uint64_t* d_mantissas, int32_t* d_exponents, int64_t* d_n_edits);

/* The edit container and statistics of cost() (src/ExtreM.cpp:272-389):
 * the 3-bit index bitmap and the snapped edits of extremz_encode, each
 * compressed to one zstd frame (the reference runs `zstd -f` on the two
 * files, src/ExtreM.cpp:312-340: level 3), plus the reference's statistics.
 * psnr / fixed_psnr are calculatePSNR (src/ExtreM.cpp:252-270) of
 * decp_before / decp_after against orig over orig's max - min.  The ratios
 * follow src/ExtreM.cpp:346-383 with original_bytes (the input file size) and
 * compressed_bytes (the lossy compressor's output, SZ3's cmpSize in the
 * reference); a ratio that needs a size the caller gave as 0 is NaN.
 * index_zst / edits_zst are nullable; their capacities must be at least
 * extremz_zstd_bound() of the raw stream sizes (ceil(3N/8) and 8 * snapped
 * edits, N at most).  Host pointers; runs on the GPU. */
typedef struct extremz_cost_stats {
  int64_t changed;          /* vertices with decp_after != decp_before */
  int64_t snapped;          /* of those, edit count 6: entries of the edit stream */
  double edit_ratio;        /* changed / N */
  double psnr, fixed_psnr;
  int64_t index_bytes, index_zst_bytes;
  int64_t edits_bytes, edits_zst_bytes;
  double ocr, cr, obr, br;
} extremz_cost_stats;

int extremz_cost(const double* orig, const double* decp_before, const double* decp_after,
                 const uint8_t* counts, int64_t n, double xi, int64_t original_bytes,
                 int64_t compressed_bytes, int zstd_level, uint8_t* index_zst,
                 int64_t index_zst_cap, uint8_t* edits_zst, int64_t edits_zst_cap,
                 extremz_cost_stats* stats);

/* zstd single frames through libzstd.so.1 (host only, no GPU).  bound: the
 * capacity a frame of raw_bytes may need.  compress returns the frame size;
 * decompress returns the content size, which exceeds cap when the buffer is
 * too small (nothing is written then).  EXTREMZ_ENOSYS without libzstd. */
int64_t extremz_zstd_bound(int64_t raw_bytes);
int64_t extremz_zstd_compress(const void* src, int64_t n, void* dst, int64_t cap, int level);
int64_t extremz_zstd_decompress(const void* src, int64_t n, void* dst, int64_t cap);

/* Split-tree check: compute_MergeT (src/ExtreM.cpp:2930-2952; persistence
 * pairs :2132-2249, merge tree and branches :1863-1918) of one float64 field
 * with its own types, labels and value order.  Per maximum in ascending
 * (value, id) order: its vertex, the saddle vertex of its persistence pair
 * (the global minimum for the global maximum), its merge-tree parent's
 * vertex, and its branch's vertices in descending (value, id) order,
 * flattened with branch_off[0..n].  Returns the number of maxima, or
 * -(needed size) - 1 when cap_max / cap_branch is too small.  Host pointers;
 * runs on the GPU. */
int extremz_split_tree(const double* field, int W, int H, int D, int32_t* max_v,
                       int32_t* pair_v, int32_t* parent_v, int32_t* branch_off,
                       int32_t* branch_v, int cap_max, int cap_branch, int* rounds);

/* The branch comparison of main.cpp:480-504 between two fields (e.g. orig
 * and the corrected field): branches and pairs are compared rank by rank;
 * maxima present in one field only count as mismatches. */
typedef struct extremz_tree_cmp {
  int64_t n_max_a, n_max_b;
  int64_t rounds_a, rounds_b;   /* pairing rounds of constructPersistencePairs */
  int64_t pair_mismatch;        /* ranks whose maximum or pair saddle differs */
  int64_t branch_mismatch;      /* ranks whose branch vertex list differs */
} extremz_tree_cmp;

int extremz_compare_split_trees(const double* a, const double* b, int W, int H, int D,
                                extremz_tree_cmp* out);

/* ---------------------------------------------------------------------------
 * Batching and serving.  No reference counterpart: the reference corrects one
 * field per process (its state is global, include/api/ExtreM.h:17-41) and its
 * GPU binary one field per run (src/ExtreM_cuda.cu:4954-6078).
 *
 * A pool owns `lanes` contexts (one HIP stream and one set of work buffers
 * each) on one device and a host worker per lane.  Jobs are admitted as a
 * lane and device memory free up: a job is admitted when the pool's device
 * footprint with the job's working set (an estimate, reconciled with the
 * buffers actually held after each job) stays within hbm_budget bytes.  Jobs
 * that do not fit are passed over by later jobs that do, unless they have
 * waited longer than `lanes` admissions; a job larger than the whole budget is
 * rejected with EXTREMZ_ENOMEM.  Each job's result is the single-field
 * result, byte for byte, whatever runs beside it.
 * ------------------------------------------------------------------------- */
typedef struct extremz_job {
  const void* orig;   /* float32 or float64 (orig_dtype) */
  double* decp;       /* corrected in place */
  uint8_t* counts;    /* nullable: receives delta_counter as uint8 */
  int W, H, D;
  int orig_dtype;     /* EXTREMZ_DTYPE_F32 / _F64 */
  int on_host;        /* 1: host buffers (staged through the lane), 0: device buffers */
  int status;         /* out: 0 or a negative EXTREMZ_E* code */
  double xi;          /* absolute bound */
  int lane;           /* out: the lane that ran the job */
  int reserved;
  double wait_ms;     /* out: submission -> admission */
  double run_ms;      /* out: admission -> completion */
  extremz_stats stats; /* out */
} extremz_job;

typedef struct extremz_pool extremz_pool;

typedef struct extremz_pool_info {
  int64_t budget_bytes;
  int64_t footprint_bytes;       /* device bytes charged to the lanes now */
  int64_t peak_footprint_bytes;  /* largest footprint so far */
  int64_t submitted, completed, running, queued;
  int lanes;
  int reserved;
} extremz_pool_info;

/* lanes in [1, 64]; hbm_budget 0 = 80% of the device's memory; opts
 * (nullable) apply to every job. */
int extremz_pool_create(int device, int lanes, int64_t hbm_budget, const extremz_opts* opts,
                        extremz_pool** out);
/* Waits for every submitted job, then frees the pool. */
int extremz_pool_destroy(extremz_pool* pool);
int extremz_pool_info_get(extremz_pool* pool, extremz_pool_info* info);

/* Batch: corrects jobs[0..n) over the lanes (largest first) and returns when
 * all are done: 0, or the first failing job's status (every job's own status
 * is in jobs[i].status). */
int extremz_batch_correct(extremz_pool* pool, extremz_job* jobs, int n);

/* Serving: queues one job and returns at once with its ticket.  The job and
 * its buffers must stay valid until the job is waited for. */
int extremz_serve_submit(extremz_pool* pool, extremz_job* job, int64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for the job of ticket; returns its
 * status, or EXTREMZ_EAGAIN if it is still queued or running.  A ticket can
 * be waited for until it has returned a status once; several threads may
 * wait for one ticket (or for it and any ticket): one of them receives the
 * status, the others EXTREMZ_EINVAL. */
int extremz_serve_wait(extremz_pool* pool, int64_t ticket, int timeout_ms);
/* Waits up to timeout_ms for any finished job that has not been waited for;
 * stores its ticket.  Returns its status or EXTREMZ_EAGAIN. */
int extremz_serve_wait_any(extremz_pool* pool, int timeout_ms, int64_t* ticket);

/* Last error message of the calling thread ("" if none). */
const char* extremz_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* EXTREMZ_H */
