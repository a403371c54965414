// engine.h -- device state shared by the kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/extremz.h"

namespace ez {

constexpr int kNbr = 14;
constexpr int kNone = -1;
// Slot pairs of one 256-side chunk of a saddle launch: at most one per
// distinct root of each side's star.
constexpr int kChunkPairs = 256 * kNbr;

// Device counters (one int array, read back with one copy).
enum Ctr {
  C_TOUCHED = 0,  // touched list length (vertices with a pending delta)
  C_DIRTY,        // dirty list length (decp changed since the last check)
  C_FRONT_A,      // frontier buffers
  C_FRONT_B,
  C_CHANGED,      // pointer jumping moved something
  C_ACT_A,        // greedy active lists
  C_ACT_B,
  C_NFC,          // number_of_false_cases
  C_NFC1,         // number_of_false_cases1
  C_WMC,          // wrong_max_counter
  C_WMC2,         // wrong_max_counter_2
  C_WMINC,        // wrong_min_counter
  C_WMINC2,       // wrong_min_counter_2
  C_WSC,          // wrong_saddle_counter
  C_WSCJ,         // wrong_saddle_counter_join
  C_GMIN,         // original global min / max (pre-snap)
  C_GMAX,
  C_DGMAX,        // global max / min of the CP lists under decp
  C_DGMIN,
  C_N2,           // list lengths
  C_N1,
  C_NMAX,
  C_NMIN,
  C_ANY_CHANGE,   // decp changed since the last CP computation
  C_FAIL,
  C_EDITED,
  C_ACT0,         // wrong-neighbour candidates of the last CP pass
  C_PCHUNK,       // chunk regions of slot pairs written by the last saddle launch
  C_NLONG,        // slots with long adjacency ranges
  C_NCHUNK,       // chunks of those ranges
  C_WN_BAIL,      // the one-block reservation solver left the list to the rounds
  C_LCHG,         // decp changed since the last CP pass of decp (list lchg)
  C_EXTD,         // ... at a listed extremum
  C_AL,           // incremental pass lists: lower star changed,
  C_CU,           // up pointer changed,
  C_CD,           // down pointer changed,
  C_QU,           // up root changed,
  C_QD,           // down root changed,
  C_DSL2,         // 2-saddle / 1-saddle sides to recompute,
  C_DSL1,
  C_NITEM,        // chunks of the extremum->saddle index to queue
  C_QLU,          // current level [C_QLU, C_QHU) of the up-root propagation queue,
  C_QHU,
  C_QLD,          // ... of the down-root one
  C_QHD,
  C_SWEEP,        // blocks of a propagation sweep (or of the walk) that finished
                  // (C_LCHG..C_SWEEP are cleared at the end of every CP pass of decp)
  C_IBAIL,        // an incremental pass overran a list (never expected; fails the call)
  C_DSLOT,        // slots to gather again
  C_SLC,          // saddle_loops calls with edits (run behind a device-side gate)
  C_PPREV,        // slots the last CP pass of decp pushed to (list pprev)
  C_PNEXT,        // slots this pass pushes to (list pnext)
  C_CL2,          // changed sides of the last CP pass
  C_CL1,
  C_CLN2,         // changed sides of this CP pass
  C_CLN1,
  C_WL2,          // sides whose decp triplet ends may differ from the original ones
  C_WL1,
  C_TOUCHED2,     // second touched / dirty list lengths: a list's consumer clears the
  C_DIRTY2,       // other one, which the next producer then fills (St::tctr / dctr)
  C_COUNT_
};

struct Geo {
  int W, H, D, WH, N;
  int off[kNbr];  // index offset of each stencil direction
  // 1: the plane adjacency of the reference's CPU path (maxNeighbors = 6 over
  // its direction table, so +-x and +-y on a depth-1 field), boundary classes
  // kPlanarCls + (x state) + 4 (y state); 0: the 14-neighbour stencil
  int planar;
};
constexpr int kPlanarCls = 64;   // first boundary class of the plane adjacency
constexpr int kLabelTileZ = 8;   // planes of a label tile (64 x 4 x kLabelTileZ voxels)
constexpr int kNumCls = 80;      // boundary classes (64 stencil + 16 plane)

// Everything a kernel needs, passed by value.
struct St {
  Geo g;
  double xi;
  int step_q;           // 0: midpoint / snap rules; q > 0: fixed step `step`, threshold q
  int f32;              // 1: the float32 fork's rule set (fixed step `step` = -xi/4 and the
                        //    star fixes in float32, applyDeltaBuffer's cap 6 -> 7)
  double step;          // -xi / q, or the float32 -xi / 4
  double thr;           // > 0: 2-saddle triplets keep the maxima at least thr above or
                        //    below the saddle (findAscPaths, src/ExtreM_cuda.cu:1878)
  long long sent;       // ordered key of the -4*xi sentinel
  const uint8_t* lut;
  const double* orig;
  double* decp;
  long long* delta;     // ordered-int64 delta buffer (d_deltaBuffer)
  uint8_t* cnt;         // delta_counter
  uint8_t* otype;       // vertex_type
  uint16_t* olow;       // original lower-star masks
  uint16_t* dlow;       // decp lower-star masks (written by the label pass)
  int *oDS, *oAS, *dDS, *dAS;
  int* stamp;
  uint8_t* dflag;
  int *touched, *dirty;
  int* ctr;
  int tctr, dctr;       // the touched / dirty list lengths in use (C_TOUCHED or C_TOUCHED2, ...)
  // CP lists (original types; equal to the decp types wherever used)
  int *S2, *S1, *MX, *MN;
  int n2, n1, nmax, nmin;
  int *of2, *ol2, *of1, *ol1;     // original triplet first / last (per list entry)
  int *df2, *dl2, *df1, *dl1;     // decp triplet first / last (per list entry)
  int *olarg, *osmall, *dlarg, *dsmall;  // per extremum, at its list position
  int* ext_id;                    // 2*position in MX (maxima) / 2*position+1 in MN, else -1
  int2* act0;                     // {2*saddle + side, candidate mask}
  long long *pv2, *pv1;           // per list position: key for the slot gather
  int *csr, *csr_off, *csr_cur;   // original adjacent saddles (list positions) per slot index
  int2* longs;                    // {slot index, first chunk} of long adjacency ranges
  int4* chunks;                   // {slot index, begin, end, -}
  uint8_t *wn2, *wn1;             // wrong_neighbors(_ds) flags
  int *wnl2, *wnl1;               // wrong_neighbors(_ds)_index
  int *minsad2, *minsad1;         // greedy reservations
  int *win_a, *win_b;             // first-saddle-wins reservations, one per extremum slot
  // Incremental CP passes of decp: what changed since the last one, and the
  // per-pass work lists.
  uint8_t* pdir;                  // steepest (up + 1) | (down + 1) << 4 per voxel; 0: itself
  uint8_t* opdir;                 // the same for the original field
  uint8_t* lflag;                 // decp changed since the last CP pass of decp
  int* lchg;                      // those vertices (C_LCHG)
  uint8_t* cflag;                 // 1: up pointer changed in this pass, 2: down pointer
  uint8_t *tu, *td;               // old up / down roots of the changed pointers
  int *spos2, *spos1;             // list position of a vertex in S2 / S1, else -1
  int *al, *cu, *cd, *ou, *od;    // pass lists (capacity icap)
  int *qu, *qd;                   // vertices whose up / down root changed (capacity N)
  int icap;
  int* sstamp;                    // per side: S2 positions, then n2 + S1 positions
  int *dsl2, *dsl1;               // sides to recompute (positions)
  int *cl2, *cl1, *cln2, *cln1;   // changed sides of the last / this pass (positions)
  int* slstamp;                   // per slot index
  int* dslot;                     // slots to gather again
  int* pstamp;                    // per slot index: pushed in the pass of this stamp
  int *pprev, *pnext;             // slots pushed by the last / this pass
  int4* citems;                   // {k0, k1, maxima?, -}: index chunks whose saddles are queued
  uint8_t* wflag;                 // per side: listed in wl2 / wl1
  int *wl2, *wl1;                 // sides whose triplet ends differed when last computed
  int iepoch;                     // stamp of this pass
};

// Modes of the saddle pass.  kModeOrig: original field, counts and records
// each saddle's (extremum, saddle) adjacencies (k_csr_fill then builds the
// per-extremum index) and leaves every saddle's key for the gather
// (k_slot_pull); kModeDec: decp, only saddles whose root set differs from the
// original one push (the others are gathered).
constexpr int kModeOrig = 0, kModeDec = 2;

struct HipError : std::runtime_error {
  explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

inline void hip_check(hipError_t e, const char* what = "hip") {
  if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}

// Host engine (engine.hip), shared with the lane pool (pool.hip).
extern thread_local std::string g_err;
extremz_ctx* new_ctx(int device);            // nullptr: no gfx950 device with that ordinal
void free_ctx(extremz_ctx* c);
size_t ctx_bytes(extremz_ctx* c);            // device bytes held by the work buffers
void ctx_release(extremz_ctx* c);            // frees the work buffers (kept: stream, LUT)
bool valid_shape(int W, int H, int D);
bool valid_opts(const extremz_opts* o);      // null, or a fixed-step q in [0, 254]
std::mutex& cache_mutex();                   // guards the per-device contexts of the host entry points
extremz_ctx* cached_ctx(int device);         // the shared context of a device (caller holds the mutex)
int ctx_device(extremz_ctx* c);
hipStream_t ctx_stream(extremz_ctx* c);
int run_correct(extremz_ctx* c, const double* d_orig, double* d_decp, int W, int H, int D,
                double xi, const extremz_opts* opts, extremz_stats* stats, uint8_t* d_counts);
// One correction with orig float32 / float64 and host or device buffers:
// host buffers are staged through the context's own device buffers.
int correct_staged(extremz_ctx* c, const void* orig, int orig_dtype, bool on_host, double* decp,
                   uint8_t* counts, int W, int H, int D, double xi, const extremz_opts* opts,
                   extremz_stats* stats);

// Split tree of one field (compute_MergeT): per maximum in ascending
// (value, id) order.
struct TreeOut {
  std::vector<int> max_v, pair_v, parent_v;
  std::vector<std::vector<int>> branch;
  int rounds = 0;
};

// Single passes over one device field on a context's stream (planar: the
// plane adjacency of Geo::planar), synchronised before they return.
void field_classify(extremz_ctx* c, int planar, const double* d_f, int W, int H, int D,
                    uint8_t* d_types, uint16_t* d_lower);
void field_labels(extremz_ctx* c, int planar, const double* d_f, int W, int H, int D, int* d_ds,
                  int* d_as);
void field_split_tree(extremz_ctx* c, int planar, const double* d_f, int W, int H, int D,
                      TreeOut& out);

// The correction of run_correct one step at a time, for a caller that drives
// the reference's step functions itself (the main.cpp:65-446 sequence over
// include/api/ExtreM.h).  Every step runs on the device state the previous
// steps left; the steps throw HipError / std::range_error on failure.
struct Stepper;
Stepper* stepper_new(extremz_ctx* c, int planar, int max_c_iters);
void stepper_free(Stepper* s);
// main.cpp:65-131 on device buffers that stay the caller's for the stepper's
// life: the original field's CP structures and the pre-snap of decp.
void stepper_begin(Stepper* s, const double* d_orig, double* d_decp, uint8_t* d_counts, int W,
                   int H, int D, double xi);
void stepper_c_loops(Stepper* s);            // c_loops
void stepper_cp(Stepper* s);                 // manifolds, computeCP and saddle maps of decp
                                             // (nothing when decp is unchanged since the last)
// The wrong-neighbour candidates of the last CP pass: {2 * saddle + side
// (1 for a 1-saddle), mask of the stencil directions whose neighbour's decp
// label differs from the original one at the same upper / lower position}.
void stepper_candidates(Stepper* s, std::vector<int2>& out);
// r_loops over the caller's wrong-neighbour lists (host arrays).
void stepper_r_loops(Stepper* s, const int* l2, int n2, const int* l1, int n1);
void stepper_s_loops(Stepper* s, int side, int do_fix);        // s_loops(_join) / get_wrong_index_*
void stepper_saddle_loops(Stepper* s, int side, int do_fix);   // saddle_loops(_join) / get_wrong_index_saddles*
// wrong_max_counter, wrong_max_counter_2, wrong_min_counter,
// wrong_min_counter_2, wrong_saddle_counter, wrong_saddle_counter_join
void stepper_counters(Stepper* s, int out[6]);
// decp was changed by the caller: the next passes start from scratch.
void stepper_decp_replaced(Stepper* s);
const uint8_t* stepper_types(Stepper* s);   // the original field's types (device)
// decp's root labels of the last CP pass (device): steepest ascent, descent
void stepper_labels(Stepper* s, const int** d_ds, const int** d_as);

// Runs fn with C++ errors translated into EXTREMZ_E* codes (message in g_err).
template <class F>
int guarded(F&& fn) {
  g_err.clear();
  try {
    return fn();
  } catch (const std::bad_alloc&) {
    g_err = "device allocation failed";
    return EXTREMZ_ENOMEM;
  } catch (const std::range_error& e) {
    g_err = std::string("iteration cap reached in ") + e.what();
    return EXTREMZ_ENOTCONV;
  } catch (const HipError& e) {
    g_err = e.what();
    return EXTREMZ_EHIP;
  } catch (const std::exception& e) {
    g_err = e.what();
    return EXTREMZ_EHIP;
  }
}

void upload_constants(const int* dx, const int* dy, const int* dz, const int* cls_key,
                      const int* lut_off);

// Kernels (kernels.hip).
__global__ void k_fill_i64(long long* p, long long v, long long n);
__global__ void k_fill_i32(int* p, int v, long long n);
__global__ void k_ctr_set(int* ctr, int idx, int v);
__global__ void k_ctr_set2(int* ctr, int a, int b, int v);
__global__ void k_ctr_export(const int* ctr, int* host, int n, int seq);
__global__ void k_f32_to_f64(const float* in, double* out, long long n);
__global__ void k_check_finite(const double* a, long long n, int* ctr, int f32);
__global__ void k_classify(Geo g, const uint8_t* lut, const double* f, uint8_t* type,
                           uint16_t* lower);
// rows per k_scan_propose block (64 x kScanTY threads)
constexpr int kScanTY = 8;
// TY rows per block: 4 (64 x 4 threads)
template <int MODE, bool FAST, int TY>
__global__ void k_mask_cols(St s, const double* f, uint8_t* type, uint16_t* lower, int* front,
                            int fctr, int nzc, int zlen);
template <bool FAST, int TY>
__global__ void k_scan_propose(St s, int* front, int fctr, int nzc, int zlen);
__global__ void k_check_list(St s, const int* list, int lctr, int with_nbrs, int* front,
                             int fctr, int epoch, int clear_dirty, int zero_idx);
__global__ void k_clear_dirty(St s, int zero_idx);
__global__ void k_check_two(St s, const int* flist, int nf, int* front, int fctr, int epoch,
                            int zero_idx);
__global__ void k_propose_c(St s, const int* front, int fctr);
__global__ void k_apply(St s, int zero_idx);
// The pre-snap a type-writing label pass of the original field makes on the
// way: decp := orig - xi (in float32 under the float32 rule set) at every
// saddle and at the global minimum and maximum (ctr[C_GMIN], ctr[C_GMAX]).
// decp == nullptr: none.
struct Presnap {
  double* decp;
  double xi;
  int f32;
  const int* ctr;
};
template <bool FAST, bool TYPE>
__global__ void k_labels_tile(Geo g, const double* f, int* ds, int* as, uint16_t* lower,
                              uint8_t* pdir, const uint8_t* lut, uint8_t* type, Presnap ps);
__global__ void k_chase(int n, int* ds, int* as);
__global__ void k_list_fill(const int* list, int n, int* arr, int val);
template <int MODE>
__global__ void k_saddle_list(St s, const double* f, const int* L2, const int* L1,
                              const uint16_t* flow, int* F2, int* La2, int* F1, int* La1,
                              int2* pairs, int* pcnt, long long* kmax, long long* kmin, int nb2,
                              int nb);
__global__ void k_slot_ids(St s, const int2* pairs, const int* pcnt, const double* f,
                           const long long* kmax, const long long* kmin, int* slot_max,
                           int* slot_min);
__global__ void k_ext_index(St s);
__global__ void k_slot_pull(St s, long long* pkey, int* pid);
__global__ void k_csr_fill(St s, const int2* pairs, const int* pcnt, int nb);
__global__ void k_slot_pull_long(St s, long long* ckey, int* cpos);
__global__ void k_slot_pull_join(St s, const long long* ckey, const int* cpos, long long* pkey,
                                 int* pid);
__global__ void k_long_slots(St s);
__global__ void k_slot_merge(St s, long long* kmax, long long* kmin, int* slot_max,
                             int* slot_min, const long long* pkey, const int* pid);
__global__ void k_scan_single(int* a, int n);
__global__ void k_slot_init(St s, long long* kmax, int* smax, long long* kmin, int* smin);
__global__ void k_list_fill64(const int* list, int n, long long* arr, long long val);
__global__ void k_ext_reduce(const int* inmax, int nmax, const int* inmin, int nmin,
                             const double* f, int* outmax, int* outmin);
__global__ void k_inc_ptrs(St s, int epoch);
__global__ void k_inc_taint(St s);
__global__ void k_inc_walk(St s);
__global__ void k_inc_sweep(St s);
__global__ void k_inc_prop(St s, int qcap);
__global__ void k_inc_mark(St s, long long* kmax, long long* kmin, int all_slots);
__global__ void k_inc_merge(St s, long long* kmax, long long* kmin, const long long* pkey,
                            const int* pid, const int2* pairs, const int* pcnt, int all_slots);
__global__ void k_inc_mark_csr(St s);
__global__ void k_inc_saddles(St s, const double* f, int2* pairs, int* pcnt, long long* kmax,
                              long long* kmin);
__global__ void k_inc_pull(St s, long long* pkey, int* pid);
__global__ void k_inc_clear(St s);
__global__ void k_inc_end_ctr(int* ctr);
__global__ void k_inc_one(St s, int epoch, long long* kmax, long long* kmin, int2* pairs,
                          int* pcnt, long long* pkey, int* pid, int end);
__global__ void k_diff_count(const void* a, const void* b, long long nbytes, int* out);
__global__ void k_act_copy(St s, int2* dst, int dctr);
__global__ void k_wn_register(St s, int2* active, int actr, int nctr);
__global__ void k_wn_commit(St s, const int2* active, int actr, int2* next, int nctr);
__global__ void k_wn_reset(St s);
__global__ void k_wn_solve(St s, int2* bufA, int2* bufB, int limit);
__global__ void k_fixpath(St s, int fast);
__global__ void k_wrong_index(St s, int side, int phase, int do_fix);
__global__ void k_wrong_index_one(St s, int side, int do_fix, int wsc);
__global__ void k_wrong_saddle(St s, int side, int do_fix, int gate);
__global__ void k_saddle_gate(int* ctr, int gate, int wsc);
template <bool TILED>
__global__ void k_lists_count(St s, int* cnt, int nb, int ntx, int ngx);
__global__ void k_lists_totals(St s, const int* cnt, int nb);
template <bool TILED>
__global__ void k_lists_write(St s, const int* cnt, int nb, int ntx, int ngx, int index);
__global__ void k_argext(const double* f, long long n, int* bmin, int* bmax, int* ctr, int check,
                         int f32);
__global__ void k_count_edited(const uint8_t* cnt, long long n, int* ctr);
__global__ void k_bitmap3(const double* before, const double* after, const uint8_t* cnt,
                          long long n, uint8_t* out, long long nbytes);
__global__ void k_edit_flags(const double* before, const double* after, const uint8_t* cnt,
                             long long n, int* flag);
__global__ void k_edit_write(const double* orig, const double* before, const int* pos,
                             const int* flag, long long n, double xi, double* edits,
                             unsigned long long* mant, int* expo);
__global__ void k_scan_block(const int* in, int* out, long long n, int* block_sums);
__global__ void k_scan_add(int* out, long long n, const int* offs);
__global__ void k_tree_triplets(Geo g, const double* f, const int* S2, int n2, const int* ds,
                                const int* ext_id, const long long* mkey, const int* MX, int* T,
                                int* cnt);
__global__ void k_tree_best(const int* T, const int* cnt, const int* S2, int n2,
                            const long long* skey, int gmax, int pass, long long* bestkey,
                            long long* bestpos);
__global__ void k_tree_pair(const int* T, const int* cnt, int nmax, const long long* bestpos,
                            int* pair_pos, int* parent, int* ptr, int* changed);
__global__ void k_tree_compress(int n, int* ptr);
__global__ void k_tree_relabel(int* T, int* cnt, int n2, const int* ptr, const long long* mkey,
                               const int* MX);
__global__ void k_list_keys(const int* list, int n, const double* f, long long* key);

}  // namespace ez
