// engine.hip -- host side of the correction engine and its C ABI.
//
// The control flow restates main.cpp:65-446 step by step; every step is a
// sequence of kernel launches on the context stream.  The host reads the
// small counter array back only where main.cpp branches on a counter.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/extremz.h"
#include "engine.h"

namespace ez {

// ------------------------------------------------------------------ LUT ----
// 14-neighbour Freudenthal stencil, src/ExtreM_cuda.cu:155-162.
static const int kDx[kNbr] = {1, -1, 0, 0, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1};
static const int kDy[kNbr] = {0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1};
static const int kDz[kNbr] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1};
// Segment order of LUT.bin, src/ExtreM_cuda.cu:84-91.
static const int kKeys[27] = {277,  153,   1050,  2085,  102,   554,  413,  1179, 2357,
                              1594, 2151,  622,   5462,  10921, 5463, 5470, 5494, 10937,
                              10925, 10923, 5599, 7543,  11197, 6014, 11963, 10991, 16383};

static bool stencil_has(int dx, int dy, int dz) {
  for (int d = 0; d < kNbr; d++)
    if (kDx[d] == dx && kDy[d] == dy && kDz[d] == dz) return true;
  return false;
}

// Number of lower / upper link components for one neighbour-existence key and
// one upper config, then the type rule of src/ExtreM_cuda.cu:458-464.
static int link_component_type(int key, int cfg) {
  int dir[kNbr], m = 0;
  for (int d = 0; d < kNbr; d++)
    if ((key >> d) & 1) dir[m++] = d;
  int comp[kNbr];
  for (int k = 0; k < m; k++) comp[k] = k;
  bool again = true;
  while (again) {  // label propagation to the smallest index
    again = false;
    for (int a = 0; a < m; a++)
      for (int b = 0; b < m; b++) {
        if (a == b || (((cfg >> a) ^ (cfg >> b)) & 1)) continue;
        int da = dir[a], db = dir[b];
        if (!stencil_has(kDx[db] - kDx[da], kDy[db] - kDy[da], kDz[db] - kDz[da])) continue;
        if (comp[b] < comp[a]) {
          comp[a] = comp[b];
          again = true;
        }
      }
  }
  int lower = 0, upper = 0;
  for (int k = 0; k < m; k++)
    if (comp[k] == k) (((cfg >> k) & 1) ? upper : lower)++;
  if (lower == 0 && upper == 1) return 0;
  if (lower == 1 && upper == 0) return 4;
  if (lower == 1 && upper == 1) return 5;
  if (lower > 1 && upper > 1) return 3;
  if (lower > 1) return 1;
  if (upper > 1) return 2;
  return 5;
}

struct LutTable {
  std::vector<uint8_t> table;
  int ref_entries = 0;
  int cls_key[kNumCls], cls_off[kNumCls];
  LutTable() {
    std::vector<int> key_off(1 << kNbr, -1);
    auto add = [&](int key) {
      key_off[key] = (int)table.size();
      int m = __builtin_popcount(key);
      for (int c = 0; c < (1 << m); c++) table.push_back((uint8_t)link_component_type(key, c));
    };
    for (int k : kKeys) add(k);
    ref_entries = (int)table.size();
    for (int cls = 0; cls < 64; cls++) {
      int st[3] = {cls & 3, (cls >> 2) & 3, (cls >> 4) & 3};
      const int* comp[3] = {kDx, kDy, kDz};
      int key = 0;
      for (int d = 0; d < kNbr; d++) {
        bool ok = true;
        for (int a = 0; a < 3; a++) {
          int c = comp[a][d];
          if (c < 0 && !(st[a] == 0 || st[a] == 2)) ok = false;
          if (c > 0 && !(st[a] == 0 || st[a] == 1)) ok = false;
        }
        if (ok) key |= 1 << d;
      }
      if (key_off[key] < 0) add(key);  // axes of length 1: extension classes
      cls_key[cls] = key;
      cls_off[cls] = key_off[key];
    }
    // the plane adjacency of the CPU path (computeAdjacency with
    // maxNeighbors = 6, src/ExtreM.cpp:543-575, on a depth-1 field): the
    // existing +-x / +-y directions of the depth-1 class.  Its link rule,
    // two neighbours joined when they share a triangle with the vertex
    // (classifyVertex1, src/ExtreM.cpp:760-818, over the triangles of
    // src/ExtreM.cpp:597-627), joins +x with +y and -x with -y: the
    // stencil's own rule restricted to these four directions.
    for (int c = 0; c < kNumCls - kPlanarCls; c++) {
      const int key = cls_key[c + 48] & 0xF;
      if (key_off[key] < 0) add(key);
      cls_key[kPlanarCls + c] = key;
      cls_off[kPlanarCls + c] = key_off[key];
    }
  }
};

static const LutTable& lut_table() {
  static LutTable t;
  return t;
}

// --------------------------------------------------------------- context ----
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void ensure(size_t n) {
    if (n <= cap) return;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      throw std::bad_alloc();
    }
    cap = n;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  size_t bytes() const { return cap * sizeof(T); }
};

struct EventPair {
  hipEvent_t a, b;
  int fam;
};

}  // namespace ez

struct extremz_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  ez::DevBuf<uint8_t> lut;
  ez::DevBuf<long long> delta, kmax, kmin;
  ez::DevBuf<uint8_t> cnt, otype, dflag, wn2, wn1;
  ez::DevBuf<uint16_t> olow, dlow;
  ez::DevBuf<int> oDS, oAS, dDS, dAS, stamp, touched, dirty, frontA, frontB;
  ez::DevBuf<int> S2, S1, MX, MN, olarg, osmall, dlarg, dsmall, minsad2, minsad1, win_a, win_b,
      lcnt, csr, csr_off, csr_cur, pid, cpos;
  ez::DevBuf<long long> pkey, ckey, pv2, pv1;
  ez::DevBuf<int2> longs;
  ez::DevBuf<int4> chunks, citems;
  ez::DevBuf<int> ext_id;
  ez::DevBuf<int> of2, ol2, of1, ol1, df2, dl2, df1, dl1, wnl2, wnl1, bmax2, bmin2;
  ez::DevBuf<int2> act0, actA, actB, pairs;
  ez::DevBuf<int> ctr;
  ez::DevBuf<double> orig_tmp, decp_tmp;
  ez::DevBuf<uint8_t> counts_tmp;
  // incremental CP passes of decp
  ez::DevBuf<uint8_t> pdir, opdir, lflag, cflag, tu, td, wflag;
  ez::DevBuf<int> lchg, spos2, spos1, ilist, qu, qd, sstamp, slists, slstamp, dslot, pcnt,
      pstamp, plists;
  ez::DevBuf<uint8_t> vsnap;  // snapshots of the self-check (EXTREMZ_VERIFY_INC)
  // the device counters as the last read left them, in coherent pinned
  // memory the export kernel writes; h_ctr[C_COUNT_] is that write's sequence
  // number
  int* h_ctr = nullptr;
  int exp_seq = 0;
  std::vector<ez::EventPair> ev_pool;
  size_t ev_used = 0;
  int timing = 0;       // bit mask of timed kernel families
  bool ev_open = false;  // the current tbegin recorded an event
  int64_t host_syncs = 0;

  // Calls f on every work buffer (not the LUT).
  template <class F>
  void each_buf(F&& f) {
    for (ez::DevBuf<int>* b :
         {&oDS, &oAS, &dDS, &dAS, &stamp, &touched, &dirty, &frontA, &frontB, &S2, &S1, &MX, &MN,
          &olarg, &osmall, &dlarg, &dsmall, &minsad2, &minsad1, &win_a, &win_b, &lcnt, &csr,
          &csr_off, &csr_cur, &pid, &cpos, &ext_id, &of2, &ol2, &of1, &ol1, &df2, &dl2, &df1,
          &dl1, &wnl2, &wnl1, &bmax2, &bmin2, &ctr, &lchg, &spos2, &spos1, &ilist, &qu, &qd,
          &sstamp, &slists, &slstamp, &dslot, &pcnt, &pstamp, &plists})
      f(*b);
    for (ez::DevBuf<long long>* b : {&delta, &kmax, &kmin, &pkey, &ckey, &pv2, &pv1}) f(*b);
    for (ez::DevBuf<uint8_t>* b :
         {&cnt, &otype, &dflag, &wn2, &wn1, &counts_tmp, &pdir, &opdir, &lflag, &cflag, &tu, &td, &wflag,
          &vsnap})
      f(*b);
    for (ez::DevBuf<uint16_t>* b : {&olow, &dlow}) f(*b);
    for (ez::DevBuf<int2>* b : {&longs, &act0, &actA, &actB, &pairs}) f(*b);
    f(chunks);
    f(citems);
    for (ez::DevBuf<double>* b : {&orig_tmp, &decp_tmp}) f(*b);
  }
};

namespace ez {

thread_local std::string g_err;

class Run {
 public:
  Run(extremz_ctx* c, const extremz_opts* o) : c_(c) {
    max_c_ = (o && o->max_c_iters > 0) ? o->max_c_iters : 100000;
    max_outer_ = (o && o->max_outer > 0) ? o->max_outer : 10000;
    c_->timing = o ? o->timing : 0;
    c_->ev_used = 0;
    step_q_ = o ? o->fixed_step_q : 0;
    planar_ = o ? o->adjacency : 0;
    f32_ = o ? o->rule_set == EXTREMZ_RULES_F32 : false;
    thr_ = o ? o->persistence : 0.0;
    const char* e = std::getenv("EXTREMZ_FULL_PASSES");
    inc_off_ = e && *e && *e != '0';
    e = std::getenv("EXTREMZ_VERIFY_INC");
    inc_verify_ = e && *e && *e != '0';
  }

  int begin(const double* d_orig, double* d_decp, int W, int H, int D, double xi,
            uint8_t* d_counts, bool presnap = true);
  int correct(const double* d_orig, double* d_decp, int W, int H, int D, double xi,
              uint8_t* d_counts);
  void classify(const double* d_f, int W, int H, int D, uint8_t* d_t, uint16_t* d_l);
  int labels(const double* d_f, int W, int H, int D, int* d_ds, int* d_as,
             uint8_t* d_t = nullptr, uint16_t* d_l = nullptr);
  void split_tree(const double* d_f, int W, int H, int D, TreeOut& out);

  int64_t counts_[EXTREMZ_STAT_COUNT_] = {};

 private:
  friend struct Stepper;
  extremz_ctx* c_;
  St s_{};
  int max_c_, max_outer_;
  int step_q_ = 0;
  int planar_ = 0;  // the CPU path's plane adjacency (Geo::planar)
  bool f32_ = false;  // the float32 fork's rule set (St::f32)
  double thr_ = 0.0;  // the persistence filter of the 2-saddle triplets (St::thr)
  int epoch_ = 0;
  bool clean_ = false;
  // the last CP pass of decp was a full one: the next incremental pass resets
  // and merges every slot (it has no list of the full pass's pushes)
  bool slots_all_ = true;
  int front_ = C_FRONT_A;  // counter index of the live frontier
  int* front_buf_ = nullptr;
  int* back_buf_ = nullptr;
  bool inc_ok_ = false;      // the decp CP state is complete (a CP pass of decp has run)
  bool inc_off_ = false;     // EXTREMZ_FULL_PASSES: every CP pass of decp is a full pass
  bool inc_verify_ = false;  // EXTREMZ_VERIFY_INC: each incremental pass is checked

  hipStream_t st() const { return c_->stream; }

  void setup_geo(int W, int H, int D) {
    s_.g.planar = planar_;
    s_.g.W = W;
    s_.g.H = H;
    s_.g.D = D;
    s_.g.WH = W * H;
    s_.g.N = W * H * D;
    for (int d = 0; d < kNbr; d++) s_.g.off[d] = kDx[d] + W * (kDy[d] + H * kDz[d]);
  }
  // 64x4 row blocks (k_classify)
  int row_blocks() const { return ((s_.g.W + 63) / 64) * ((s_.g.H + 3) / 4) * s_.g.D; }
  // 64 x ty columns cut into nzc chunks of zlen planes, about target blocks
  // (k_mask_cols: 64 x 4, k_scan_propose: 64 x kScanTY)
  struct ColGrid {
    int blocks, nzc, zlen;
  };
  ColGrid col_grid(int ty = 4, int target = kColBlocks) const {
    const int ncol = ((s_.g.W + 63) / 64) * ((s_.g.H + ty - 1) / ty);
    int nzc = std::max(1, std::min(s_.g.D, (target + ncol - 1) / ncol));
    const int zlen = (s_.g.D + nzc - 1) / nzc;
    nzc = (s_.g.D + zlen - 1) / zlen;
    return {ncol * nzc, nzc, zlen};
  }
  // one full false-critical-point scan of decp into the frontier list fbuf
  void check_full(int* fbuf, int fctr) {
    const ColGrid cg = col_grid();
    tbegin(EXTREMZ_K_CHECK_FULL);
    auto* kern = s_.g.planar ? k_mask_cols<1, false, 4> : k_mask_cols<1, true, 4>;
    kern<<<cg.blocks, dim3(64, 4, 1), 0, st()>>>(s_, s_.decp, nullptr, nullptr, fbuf, fctr,
                                                 cg.nzc, cg.zlen);
    tend();
  }
  // the same scan with the frontier's proposals made (delta and the touched
  // list)
  void scan_propose(int* fbuf, int fctr) {
    const ColGrid cg = col_grid(kScanTY, kScanBlocks);
    tbegin(EXTREMZ_K_SCAN_PROPOSE);
    auto* kern = s_.g.planar ? k_scan_propose<false, kScanTY> : k_scan_propose<true, kScanTY>;
    kern<<<cg.blocks, dim3(64, kScanTY, 1), 0, st()>>>(s_, fbuf, fctr, cg.nzc, cg.zlen);
    tend();
  }
  // 64x4xkLabelTileZ label tiles
  int tile_blocks() const {
    return ((s_.g.W + 63) / 64) * ((s_.g.H + 3) / 4) *
           ((s_.g.D + kLabelTileZ - 1) / kLabelTileZ);
  }
  // Capacity of the incremental pass's vertex lists: a pass with nch changed
  // vertices touches at most 15 * nch vertices, and runs only when that fits.
  static long long inc_cap(long long N) { return N / 4 + 1024; }
  // Largest share of changed vertices (1 / kIncDiv of the grid) taken
  // incrementally; more, and the full pass is faster.
  static constexpr long long kIncDiv = 256;
  // Whole-grid levels of the root propagation before the one-block remainder.
  static constexpr int kSweeps = 3, kSweepBlocks = 256;
  // Blocks of a saddle launch (each takes a contiguous range of 256-side
  // chunks, consecutive planes of one 32x32 column, so a block's plane
  // neighbours stay cached): at least this many, else about 4 chunks each.
  static constexpr int kSaddleBlocks = 4096;
  // Blocks of a column sweep (one frontier append each).
  static constexpr int kColBlocks = 4096;
  // k_scan_propose: 64 x kScanTY columns, about kScanBlocks blocks.  Rows of
  // 4 / 8 / 16: 2.97 / 2.70 / 2.53 ms per 512^3 scan (fewer halo words per
  // owned voxel), but a 1024-thread block needs a whole CU free: beside other
  // contexts' kernels (a serving pool) rows of 16 wait for one
  static constexpr int kScanBlocks = 1024;
  // Blocks of the incremental saddle launches (sizes known on the device only).
  static constexpr int kIncSaddleBlocks = 2048;
  // Wrong-neighbour reservation rounds launched per read of the list length.
  static constexpr int kWnRounds = 4;
  // Listed sides up to which the wrong-index phases run in one block.
  static constexpr int kWrongOne = 4096;
  // Candidate lists up to which k_wn_solve runs every reservation round in
  // one block; longer ones take the multi-block rounds.
  static constexpr int kWnOneMax = 4096;
  // Blocks of the first stage of the decp extremum reduction.
  static constexpr int kExtBlocks = 2048;
  // Changed vertices up to which an incremental pass runs in one block (its
  // lists then hold a few thousand entries at most).
  static constexpr long long kIncOne = 64;
  bool ended_ = false;  // the last CP pass cleared its own state (cp_end has no launch)
  static int blocks_for(long long n) {
    long long b = (n + 255) / 256;
    return (int)std::max<long long>(1, std::min<long long>(b, 2048));
  }

  // ---- timing
  void tbegin(int fam) {
    c_->ev_open = (c_->timing >> fam) & 1;
    if (!c_->ev_open) return;
    if (c_->ev_used == c_->ev_pool.size()) {
      EventPair p;
      ez::hip_check(hipEventCreate(&p.a), "hipEventCreate");
      ez::hip_check(hipEventCreate(&p.b), "hipEventCreate");
      c_->ev_pool.push_back(p);
    }
    EventPair& p = c_->ev_pool[c_->ev_used];
    p.fam = fam;
    ez::hip_check(hipEventRecord(p.a, st()), "hipEventRecord");
  }
  void tend() {
    if (!c_->ev_open) return;
    c_->ev_open = false;
    ez::hip_check(hipEventRecord(c_->ev_pool[c_->ev_used].b, st()), "hipEventRecord");
    c_->ev_used++;
  }

  // ---- counters
  void set_ctr(int idx, int v) { k_ctr_set<<<1, 64, 0, st()>>>(s_.ctr, idx, v); }
  void set_ctr2(int a, int b, int v) { k_ctr_set2<<<1, 64, 0, st()>>>(s_.ctr, a, b, v); }
  // Reads the device counters: a one-wavefront kernel writes them, then its
  // sequence number, into coherent pinned host memory, and the host waits on
  // that number (a stream synchronisation costs a few microseconds more per
  // read, and much more with several contexts reading at once).
  void pull() {
    ez::hip_check(hipGetLastError(), "kernel launch");
    const int seq = ++c_->exp_seq;
    k_ctr_export<<<1, 64, 0, st()>>>(s_.ctr, c_->h_ctr, C_COUNT_, seq);
    ez::hip_check(hipGetLastError(), "kernel launch");
    const volatile int* flag = c_->h_ctr + C_COUNT_;
    for (long long spin = 1; __atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq; spin++) {
      if (spin % 256) {
        // a short wait spins; past ~0.1 ms (a full-grid pass) the thread
        // gives its core to others between polls
        if (spin < 4096) __builtin_ia32_pause();
        else std::this_thread::yield();
        continue;
      }
      // every few microseconds: a failed kernel ends the wait with its error,
      // a drained stream with a value that must have arrived
      const hipError_t e = hipStreamQuery(st());
      if (e == hipSuccess) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != seq)
          throw HipError("counter export did not arrive");
        break;
      }
      if (e != hipErrorNotReady) ez::hip_check(e, "hipStreamQuery");
    }
    c_->host_syncs++;
    if (c_->h_ctr[C_IBAIL]) throw HipError("incremental CP pass overran a work list");
  }
  int h(int idx) const { return c_->h_ctr[idx]; }

  // The global minimum / maximum of f by (value, id) into ctr[C_GMIN] /
  // ctr[C_GMAX]; with check, a value that is not finite (or not a float32
  // under the float32 rules) sets ctr[C_FAIL].
  void argext(const double* f, int check) {
    c_->bmax2.ensure(kExtBlocks);
    c_->bmin2.ensure(kExtBlocks);
    const int nb = std::min(blocks_for(s_.g.N), kExtBlocks);
    k_argext<<<nb, 256, 0, st()>>>(f, s_.g.N, c_->bmin2.p, c_->bmax2.p, s_.ctr, check, s_.f32);
    k_ext_reduce<<<1, 256, 0, st()>>>(c_->bmax2.p, nb, c_->bmin2.p, nb, f, s_.ctr + C_GMAX,
                                      s_.ctr + C_GMIN);
  }
  void allocate(long long N);
  void list_compaction(int index);
  void slot_pass(int dec, const double* f, const int* L2, const int* L1, const uint16_t* flow,
                 int* F2, int* La2, int* F1, int* La1, int* smax, int* smin);
  void apply(long long bound = 1LL << 62, int also = -1);
  void dirty_consumed();
  void label_pass(const double* f, int* ds, int* as, uint16_t* lower, uint8_t* pdir = nullptr,
                  uint8_t* type = nullptr, Presnap ps = Presnap{});
  void c_loops(bool entered = false);
  void c_loops_entry();
  bool proposed_ = false;  // the frontier's proposals are made already
  void cp_dec(bool force, bool fresh = false);
  void cp_inc(long long nch);
  void cp_end();
  void cp_full();
  void verify_inc();
  void wrong_neighbors();
  void r_loops();
  bool s_loops(int side, int do_fix, int known = -1, bool gate = false);
  void saddle_loops(int side, int do_fix);
  void saddle_loops_if(int side, bool gated = false);
};

// The CP lists S2 / S1 / MX / MN of the types in s_.otype, in the column
// order of k_lists_count (32 x 32 columns for fields at least that wide).
void Run::list_compaction(int index) {
  const Geo& g = s_.g;
  const int tiled = g.W >= 32 && g.H >= 32;
  // units of 1024 traversal positions (a z-plane of a 32 x 32 column, or 1024
  // vertices), eight per block: neighbouring columns of one plane
  const int ntx = (g.W + 31) / 32, nty = (g.H + 31) / 32, ngx = (ntx + 7) / 8;
  const int nb = (int)(tiled ? (long long)ntx * nty * g.D : (g.N + 1023) / 1024);
  const int blocks = (int)(tiled ? (long long)ngx * nty * g.D : (nb + 7LL) / 8);
  // the 4 * nb unit counts (plus a zero that receives the total) scanned as
  // one array; block sums after them
  const long long n1 = 4LL * nb + 1;
  const int nblk = (int)((n1 + 1023) / 1024);
  c_->lcnt.ensure(n1 + nblk + 1);
  int* cnt = c_->lcnt.p;
  int* sums = cnt + n1;
  k_fill_i32<<<1, 64, 0, st()>>>(cnt + 4LL * nb, 0, 1);
  auto* count = tiled ? k_lists_count<true> : k_lists_count<false>;
  auto* write = tiled ? k_lists_write<true> : k_lists_write<false>;
  count<<<blocks, 256, 0, st()>>>(s_, cnt, nb, ntx, ngx);
  k_scan_block<<<nblk, 1024, 0, st()>>>(cnt, cnt, n1, sums);
  k_scan_single<<<1, 1024, 0, st()>>>(sums, nblk);
  k_scan_add<<<nblk, 1024, 0, st()>>>(cnt, n1, sums);
  k_lists_totals<<<1, 64, 0, st()>>>(s_, cnt, nb);
  write<<<blocks, 256, 0, st()>>>(s_, cnt, nb, ntx, ngx, index);
}

void Run::allocate(long long N) {
  extremz_ctx* c = c_;
  c->delta.ensure(N);
  c->kmax.ensure(N);
  c->kmin.ensure(N);
  c->ext_id.ensure(N);
  for (DevBuf<uint8_t>* b : {&c->cnt, &c->otype, &c->dflag, &c->wn2, &c->wn1}) b->ensure(N);
  c->olow.ensure(N);
  c->dlow.ensure(N);
  for (DevBuf<int>* b : {&c->oDS, &c->oAS, &c->dDS, &c->dAS, &c->stamp, &c->touched, &c->dirty,
                         &c->frontA, &c->frontB, &c->S2, &c->S1, &c->MX, &c->MN, &c->olarg,
                         &c->osmall, &c->dlarg, &c->dsmall, &c->minsad2, &c->minsad1,
                         &c->of2, &c->ol2, &c->of1, &c->ol1, &c->df2,
                         &c->dl2, &c->df1, &c->dl1, &c->wnl2, &c->wnl1})
    b->ensure(N);
  for (DevBuf<uint8_t>* b : {&c->pdir, &c->opdir, &c->lflag, &c->cflag, &c->tu, &c->td})
    b->ensure(N);
  for (DevBuf<int>* b : {&c->lchg, &c->spos2, &c->spos1, &c->qu, &c->qd}) b->ensure(N);
  c->ilist.ensure(5 * inc_cap(N));
  c->bmax2.ensure(kExtBlocks);
  c->bmin2.ensure(kExtBlocks);
  c->act0.ensure(2 * N);
  c->actA.ensure(2 * N);
  c->actB.ensure(2 * N);
  c->ctr.ensure(C_COUNT_);
}

// computeCP + triplets + largest/smallest saddle maps of field f in one
// full-grid pass (plus the slot tie-break pass); with dec, also the
// wrong-neighbour candidates and decp's global maximum / minimum.
void Run::slot_pass(int dec, const double* f, const int* L2, const int* L1, const uint16_t* flow,
                    int* F2, int* La2, int* F1, int* La1, int* smax, int* smin) {
  long long* kmax = c_->kmax.p;
  long long* kmin = c_->kmin.p;
  if (dec) slots_all_ = true;
  tbegin(EXTREMZ_K_SLOT_GATHER);
  k_slot_init<<<blocks_for(std::max(s_.nmax, s_.nmin)), 256, 0, st()>>>(s_, kmax, smax, kmin,
                                                                          smin);
  tend();
  const int nb2 = (s_.n2 + 255) / 256, nb1 = (s_.n1 + 255) / 256;
  const int ns = s_.nmax + s_.nmin;
  auto saddle_pass = [&](int mode) {
    if (!(nb2 + nb1)) {
      set_ctr(C_PCHUNK, 0);
      return;
    }
    tbegin(mode == kModeOrig ? EXTREMZ_K_SADDLE_ORIG : EXTREMZ_K_SADDLE_PASS);
    auto* kern = mode == kModeOrig ? k_saddle_list<kModeOrig> : k_saddle_list<kModeDec>;
    const int nb = nb2 + nb1;
    kern<<<std::min(nb, std::max(kSaddleBlocks, nb / 4)), 256, 0, st()>>>(
        s_, f, L2, L1, flow, F2, La2, F1, La1, c_->pairs.p, c_->pcnt.p, kmax, kmin, nb2,
        nb2 + nb1);
    tend();
  };
  if (!dec) {
    // original field: the (extremum, saddle) adjacencies are indexed per
    // extremum (for the decp passes too), and each extremum gathers its
    // largest / smallest saddle over them
    k_fill_i32<<<blocks_for(ns + 1), 256, 0, st()>>>(s_.csr_off, 0, ns + 1);
    saddle_pass(kModeOrig);
    {  // exclusive scan of the ns counts (entry ns, zero, receives the total)
      const long long n1 = (long long)ns + 1;
      const int nblk = (int)((n1 + 1023) / 1024);
      c_->lcnt.ensure((size_t)nblk + 1);
      k_scan_block<<<nblk, 1024, 0, st()>>>(s_.csr_off, s_.csr_off, n1, c_->lcnt.p);
      k_scan_single<<<1, 1024, 0, st()>>>(c_->lcnt.p, nblk);
      k_scan_add<<<nblk, 1024, 0, st()>>>(s_.csr_off, n1, c_->lcnt.p);
    }
    int total = 0;
    ez::hip_check(hipMemcpyAsync(&total, s_.csr_off + ns, sizeof(int), hipMemcpyDeviceToHost,
                                 st()),
                  "hipMemcpyAsync");
    ez::hip_check(hipStreamSynchronize(st()), "hipStreamSynchronize");
    c_->host_syncs++;
    c_->csr.ensure((size_t)total + 1);
    s_.csr = c_->csr.p;
    // every long range has more than kPullShort (256) entries
    const size_t nch = (size_t)total / 256 + (size_t)total / 1024 + 2;
    c_->chunks.ensure(nch);
    c_->ckey.ensure(nch);
    c_->cpos.ensure(nch);
    // the incremental pass cuts the index ranges of changed extrema into
    // chunks of 1024: at most one per extremum plus one per 1024 entries
    c_->citems.ensure((size_t)ns + total / 1024 + 2);
    s_.citems = c_->citems.p;
    s_.chunks = c_->chunks.p;
    // after the buffers above exist: the launch captures their pointers
    set_ctr2(C_NLONG, C_NCHUNK, 0);
    k_long_slots<<<blocks_for(ns), 256, 0, st()>>>(s_);
    ez::hip_check(hipMemcpyAsync(s_.csr_cur, s_.csr_off, sizeof(int) * (ns + 1),
                                 hipMemcpyDeviceToDevice, st()),
                  "hipMemcpyAsync");
    tbegin(EXTREMZ_K_CSR_FILL);
    k_csr_fill<<<std::max(1, std::min(nb2 + nb1, 8192)), 256, 0, st()>>>(
        s_, c_->pairs.p, c_->pcnt.p, nb2 + nb1);
    tend();
    tbegin(EXTREMZ_K_SLOT_GATHER);
    k_slot_pull<<<std::max(1, (ns + 31) / 32), 256, 0, st()>>>(s_, c_->pkey.p, c_->pid.p);
    k_slot_pull_long<<<512, 256, 0, st()>>>(s_, c_->ckey.p, c_->cpos.p);
    k_slot_pull_join<<<64, 256, 0, st()>>>(s_, c_->ckey.p, c_->cpos.p, c_->pkey.p, c_->pid.p);
  } else {
    // every side is recomputed: the list of sides with differing triplet ends
    // starts empty
    hip_check(hipMemsetAsync(s_.wflag, 0, (size_t)s_.n2 + s_.n1 + 1, st()), "memset");
    set_ctr2(C_WL2, C_WL1, 0);
    saddle_pass(kModeDec);
    tbegin(EXTREMZ_K_SLOT_GATHER);
    k_slot_pull<<<std::max(1, (ns + 31) / 32), 256, 0, st()>>>(s_, c_->pkey.p, c_->pid.p);
    k_slot_pull_long<<<512, 256, 0, st()>>>(s_, c_->ckey.p, c_->cpos.p);
    k_slot_pull_join<<<64, 256, 0, st()>>>(s_, c_->ckey.p, c_->cpos.p, c_->pkey.p, c_->pid.p);
  }
  k_slot_merge<<<blocks_for((long long)s_.nmax + s_.nmin), 256, 0, st()>>>(
      s_, kmax, kmin, smax, smin, c_->pkey.p, c_->pid.p);
  // ties among the pushed saddles: one wavefront per chunk region
  if (dec)
    k_slot_ids<<<blocks_for(64LL * (nb2 + nb1)), 256, 0, st()>>>(s_, c_->pairs.p, c_->pcnt.p, f,
                                                                 kmax, kmin, smax, smin);
  if (dec) {
    // decp global maximum / minimum among the listed extrema
    // the value gathers are scattered: one short stride per lane over kExtBlocks blocks
    k_ext_reduce<<<kExtBlocks, 256, 0, st()>>>(s_.MX, s_.nmax, s_.MN, s_.nmin, f, c_->bmax2.p,
                                               c_->bmin2.p);
    k_ext_reduce<<<1, 256, 0, st()>>>(c_->bmax2.p, kExtBlocks, c_->bmin2.p, kExtBlocks, f,
                                      s_.ctr + C_DGMAX, s_.ctr + C_DGMIN);
  }
  tend();
}

// apply_delta over the touched list (at most `bound` entries), then the list
// is emptied (with the counter `also`, when given).
void Run::apply(long long bound, int also) {
  tbegin(EXTREMZ_K_APPLY);
  k_apply<<<blocks_for(std::min<long long>(bound, s_.g.N)), 256, 0, st()>>>(s_, also);
  tend();
  s_.tctr = s_.tctr == C_TOUCHED ? C_TOUCHED2 : C_TOUCHED;  // k_apply cleared it
}
void Run::dirty_consumed() { s_.dctr = s_.dctr == C_DIRTY ? C_DIRTY2 : C_DIRTY; }

// ComputeDescendingManifold + ComputeAscendingManifold on field f (D7): tile
// resolution, then a chase of whatever is left (a separate pointer-jumping
// round before the chase costs more than the chase saves).  No host round
// trip.
void Run::label_pass(const double* f, int* ds, int* as, uint16_t* lower, uint8_t* pdir,
                     uint8_t* type, Presnap ps) {
  counts_[EXTREMZ_STAT_LABEL_PASSES]++;
  tbegin(EXTREMZ_K_LABEL_TILE);
  // every voxel on the ordered scan: W >= 3 on the stencil
  const bool fast = s_.g.W >= 3 && !s_.g.planar;
  auto* kern = fast ? (type ? k_labels_tile<true, true> : k_labels_tile<true, false>)
                    : (type ? k_labels_tile<false, true> : k_labels_tile<false, false>);
  kern<<<tile_blocks(), dim3(64, 4, 1), 0, st()>>>(s_.g, f, ds, as, lower, pdir, c_->lut.p,
                                                   type, ps);
  tend();
  int nb = (s_.g.N + 255) / 256;
  tbegin(EXTREMZ_K_LABEL_CHASE);
  k_chase<<<nb, 256, 0, st()>>>(s_.g.N, ds, as);
  tend();
}

// The start of c_loops: the first false-critical-point check.
void Run::c_loops_entry() {
  counts_[EXTREMZ_STAT_C_LOOPS_CALLS]++;
  // both frontier lengths are zero here: the correction starts with zeroed
  // counters and every c_loops ends with both at zero (below)
  front_ = C_FRONT_A;
  front_buf_ = c_->frontA.p;
  back_buf_ = c_->frontB.p;
  epoch_++;
  // whether the frontier's proposals are made already.  The first scan of a
  // correction meets the compressor's false critical points, a dense
  // frontier (5-30% of the grid) whose proposals cost more as scattered
  // atomics than the scan itself: that scan proposes on the way.  Rescans
  // after many edits find few false points, and the plain scan is cheaper.
  proposed_ = false;
  if (!clean_) {
    scan_propose(front_buf_, front_);
    proposed_ = true;
    k_clear_dirty<<<blocks_for(s_.g.N), 256, 0, st()>>>(s_, -1);
  } else {
    tbegin(EXTREMZ_K_CHECK_INCR);
    k_check_list<<<blocks_for(s_.g.N), 256, 0, st()>>>(s_, s_.dirty, s_.dctr, 1, front_buf_,
                                                        front_, epoch_, 1, -1);
    tend();
  }
  dirty_consumed();
}

// c_loops, src/ExtreM.cpp:1505-1583.  The first call scans the full grid; later
// calls start from the vertices edited since the last clean state.  With
// entered, c_loops_entry has run and the last counter read came after it.
void Run::c_loops(bool entered) {
  if (!entered) c_loops_entry();
  bool proposed = proposed_;
  for (bool read = !entered;; read = true) {
    if (read) pull();
    int nf = h(front_);
    if (nf == 0) break;
    counts_[EXTREMZ_STAT_FRONTIER_ITEMS] += nf;
    if (++counts_[EXTREMZ_STAT_C_LOOP_ITERS] > max_c_) throw std::range_error("c_loops");
    if (!proposed) {
      tbegin(EXTREMZ_K_PROPOSE);
      k_propose_c<<<blocks_for(nf), 256, 0, st()>>>(s_, front_buf_, front_);
      tend();
    }
    proposed = false;
    // a frontier vertex proposes to itself or at most its 14 neighbours, so
    // at most 15 nf vertices are touched and changed
    const long long nt = 15LL * nf;
    int back = front_ == C_FRONT_A ? C_FRONT_B : C_FRONT_A;
    apply(nt, back);
    epoch_++;
    if (nt >= s_.g.N / 8) {
      // many edits: one full scan gives the same set as the re-checks below
      // (old entries that stay false plus the edited neighbourhoods), faster
      check_full(back_buf_, back);
      k_clear_dirty<<<blocks_for(nt), 256, 0, st()>>>(s_, front_);
    } else {
      // the old entries and the edited neighbourhoods; the consumed front
      // length is cleared here, so the loop ends with both lengths at zero
      tbegin(EXTREMZ_K_CHECK_INCR);
      k_check_two<<<blocks_for(nf + 15 * nt), 256, 0, st()>>>(s_, front_buf_, nf, back_buf_, back,
                                                              epoch_, front_);
      tend();
    }
    dirty_consumed();
    std::swap(front_buf_, back_buf_);
    front_ = back;
  }
  clean_ = true;
}

// Manifolds + computeCP + largest/smallest saddle maps + wrong-neighbour
// candidates of decp (main.cpp:290-304 and the comparisons of 313-323).
// Skipped when decp has not changed since the last call: everything it
// computes depends on decp only.
// With fresh, the counters read last are still current (no launch since).
void Run::cp_dec(bool force, bool fresh) {
  long long nch = 0;
  if (!force) {
    if (!fresh) pull();
    if (!h(C_ANY_CHANGE)) return;
    nch = h(C_LCHG);
  }
  // few changes since the last pass: only what they can reach is recomputed
  // (the self-check takes every pass that fits the lists, to exercise them)
  const bool inc = inc_ok_ && !inc_off_ && 15 * nch <= s_.icap &&
                   (inc_verify_ || nch <= s_.g.N / kIncDiv);
  if (inc) cp_inc(nch);
  else cp_full();
  cp_end();
  inc_ok_ = true;
}

void Run::cp_full() {
  label_pass(s_.decp, s_.dDS, s_.dAS, s_.dlow, s_.pdir);
  slot_pass(1, s_.decp, s_.dDS, s_.dAS, s_.dlow, s_.df2, s_.dl2, s_.df1, s_.dl1, s_.dlarg,
            s_.dsmall);
}

// The incremental CP pass (the kernels' section "incremental CP pass"), with
// no host read: list sizes stay on the device.  A list overrun (C_IBAIL,
// never expected: every list holds its worst case) fails the call at the
// next counter read.
void Run::cp_inc(long long nch) {
  counts_[EXTREMZ_STAT_INC_PASSES]++;
  epoch_++;
  s_.iepoch = epoch_;
  const int b1 = blocks_for(15 * nch);
  tbegin(EXTREMZ_K_INCREMENTAL);
  long long* kmax = c_->kmax.p;
  long long* kmin = c_->kmin.p;
  if (nch <= kIncOne && !slots_all_) {
    // few changes: the whole pass (and, unless it is checked next, its end)
    // in one block
    ended_ = !inc_verify_;
    k_inc_one<<<1, 1024, 0, st()>>>(s_, epoch_, kmax, kmin, c_->pairs.p, c_->pcnt.p, c_->pkey.p,
                                    c_->pid.p, ended_ ? 1 : 0);
    tend();
    if (inc_verify_) verify_inc();
    return;
  }
  k_inc_ptrs<<<b1, 256, 0, st()>>>(s_, epoch_);
  k_inc_taint<<<b1, 256, 0, st()>>>(s_);
  k_inc_walk<<<std::min(b1, 512), 256, 0, st()>>>(s_);
  // the first levels of the root propagation over the whole grid, the rest in
  // one block
  for (int l = 0; l < kSweeps; l++) k_inc_sweep<<<kSweepBlocks, 256, 0, st()>>>(s_);
  k_inc_prop<<<1, 1024, 0, st()>>>(s_, s_.g.N);
  k_inc_mark<<<1024, 256, 0, st()>>>(s_, kmax, kmin, slots_all_);
  k_inc_mark_csr<<<512, 256, 0, st()>>>(s_);
  // the queued sides, their slots and ties; sizes stay on the device
  k_inc_saddles<<<kIncSaddleBlocks, 256, 0, st()>>>(s_, s_.decp, c_->pairs.p, c_->pcnt.p, kmax,
                                                     kmin);
  k_inc_pull<<<1024, 256, 0, st()>>>(s_, c_->pkey.p, c_->pid.p);
  k_inc_merge<<<slots_all_ ? blocks_for((long long)s_.nmax + s_.nmin) : kIncSaddleBlocks, 256, 0,
                st()>>>(s_, kmax, kmin, c_->pkey.p, c_->pid.p, c_->pairs.p, c_->pcnt.p,
                        slots_all_);
  slots_all_ = false;
  k_slot_ids<<<kIncSaddleBlocks, 256, 0, st()>>>(s_, c_->pairs.p, c_->pcnt.p, s_.decp, kmax,
                                                 kmin, s_.dlarg, s_.dsmall);
  tend();
  if (inc_verify_) verify_inc();
}

// End of every CP pass of decp: the pass's flags are cleared, the changed
// sides become the last pass's, and the change list starts empty.
void Run::cp_end() {
  if (!ended_) {
    k_inc_clear<<<blocks_for(15 * (long long)s_.icap), 256, 0, st()>>>(s_);
    k_inc_end_ctr<<<1, 64, 0, st()>>>(s_.ctr);
  }
  ended_ = false;
  std::swap(s_.cl2, s_.cln2);
  std::swap(s_.cl1, s_.cln1);
  std::swap(s_.pprev, s_.pnext);
}

// Self-check of an incremental pass: its state against the full pass's on
// the same decp (which then stands).  Throws on any difference.
void Run::verify_inc() {
  struct Arr {
    const char* name;
    void* p;
    long long bytes;
  };
  const long long N = s_.g.N, ns = (long long)s_.nmax + s_.nmin;
  const Arr arrs[] = {{"dDS", s_.dDS, 4 * N},         {"dAS", s_.dAS, 4 * N},
                      {"dlow", s_.dlow, 2 * N},       {"pdir", s_.pdir, N},
                      {"df2", s_.df2, 4LL * s_.n2},   {"dl2", s_.dl2, 4LL * s_.n2},
                      {"df1", s_.df1, 4LL * s_.n1},   {"dl1", s_.dl1, 4LL * s_.n1},
                      {"pv2", s_.pv2, 8LL * s_.n2},   {"pv1", s_.pv1, 8LL * s_.n1},
                      {"dlarg", s_.dlarg, 4LL * s_.nmax}, {"dsmall", s_.dsmall, 4LL * s_.nmin},
                      {"pkey", c_->pkey.p, 8 * ns},   {"pid", c_->pid.p, 4 * ns},
                      {"kmax", c_->kmax.p, 8LL * s_.nmax}, {"kmin", c_->kmin.p, 8LL * s_.nmin}};
  const int na = (int)(sizeof(arrs) / sizeof(arrs[0]));
  long long total = 0;
  for (const Arr& a : arrs) total += a.bytes;
  c_->vsnap.ensure((size_t)total + 64);
  hipStream_t q = st();
  long long off = 0;
  for (const Arr& a : arrs) {
    if (a.bytes) hip_check(hipMemcpyAsync(c_->vsnap.p + off, a.p, a.bytes,
                                          hipMemcpyDeviceToDevice, q), "hipMemcpyAsync");
    off += a.bytes;
  }
  // (the number of recorded slot pairs depends on the order of the atomics,
  // only the slots they settle do not)
  int hc[3];
  const int keep[3] = {C_DGMAX, C_DGMIN, C_ACT0};
  pull();
  for (int k = 0; k < 3; k++) hc[k] = h(keep[k]);
  // the full pass on the same decp, over the same change lists
  label_pass(s_.decp, s_.dDS, s_.dAS, s_.dlow, s_.pdir);
  slot_pass(1, s_.decp, s_.dDS, s_.dAS, s_.dlow, s_.df2, s_.dl2, s_.df1, s_.dl1, s_.dlarg,
            s_.dsmall);
  DevBuf<int> diff;
  diff.ensure(na);
  hip_check(hipMemsetAsync(diff.p, 0, sizeof(int) * na, q), "memset");
  off = 0;
  for (int k = 0; k < na; k++) {
    if (arrs[k].bytes)
      k_diff_count<<<blocks_for(arrs[k].bytes), 256, 0, q>>>(c_->vsnap.p + off, arrs[k].p,
                                                             arrs[k].bytes, diff.p + k);
    off += arrs[k].bytes;
  }
  std::vector<int> hd(na);
  hip_check(hipMemcpyAsync(hd.data(), diff.p, sizeof(int) * na, hipMemcpyDeviceToHost, q),
            "hipMemcpyAsync");
  pull();
  std::string bad;
  for (int k = 0; k < na; k++)
    if (hd[k]) bad += std::string(" ") + arrs[k].name + "(" + std::to_string(hd[k]) + " B)";
  const char* cn[3] = {"dgmax", "dgmin", "act0"};
  for (int k = 0; k < 3; k++)
    if (hc[k] != h(keep[k])) bad += std::string(" ") + cn[k];
  if (!bad.empty()) throw HipError("incremental CP pass differs from the full pass:" + bad);
}

// get_wrong_split_neighbors / get_wrong_join_neighbors over all saddles in
// saddle-id order (main.cpp:309-323), as rounds of deterministic reservations
// on the candidates of the last CP pass.
void Run::wrong_neighbors() {
  // short candidate lists: every round in one block, one host read
  tbegin(EXTREMZ_K_RESERVE);
  k_wn_solve<<<1, 1024, 0, st()>>>(s_, c_->actA.p, c_->actB.p, kWnOneMax);
  tend();
  pull();
  if (h(C_WN_BAIL) == 2) throw HipError("wrong-neighbour reservations did not finish");
  if (!h(C_WN_BAIL)) return;
  k_wn_reset<<<blocks_for(s_.g.N), 256, 0, st()>>>(s_);
  set_ctr2(C_NFC, C_NFC1, 0);
  int2* act = c_->actA.p;
  int2* nxt = c_->actB.p;
  int actr = C_ACT_A, nctr = C_ACT_B;
  k_act_copy<<<blocks_for(2LL * s_.g.N), 256, 0, st()>>>(s_, act, actr);
  long long guard = (long long)s_.n2 + s_.n1 + 2;
  // the list length is read every kWnRounds rounds; it never grows, so the
  // grid of the last read covers the rounds until the next one (rounds on an
  // empty list do nothing)
  for (long long round = 0;;) {
    pull();
    int na = h(actr);
    if (na == 0) break;
    // each round commits at least the lowest active saddle
    if (round > guard) throw HipError("wrong-neighbour reservations did not finish");
    int b = blocks_for(na);
    tbegin(EXTREMZ_K_RESERVE);
    for (int r = 0; r < kWnRounds; r++, round++) {
      k_wn_register<<<b, 256, 0, st()>>>(s_, act, actr, nctr);
      k_wn_commit<<<b, 256, 0, st()>>>(s_, act, actr, nxt, nctr);
      std::swap(act, nxt);
      std::swap(actr, nctr);
    }
    tend();
  }
}

// r_loops, src/ExtreM.cpp:2577-2593.
// Called after a counter read with no edit since (the outer loop's first
// test, or the wrong-neighbour read of the loop's last repetition, which
// makes no edits): it holds the wrong-neighbour counts and says whether decp
// changed since its last CP pass.
void Run::r_loops() {
  counts_[EXTREMZ_STAT_R_LOOP_CALLS]++;
  const int fast = inc_ok_ && !h(C_ANY_CHANGE);
  tbegin(EXTREMZ_K_PROPOSE);
  k_fixpath<<<blocks_for(std::max(1, h(C_NFC) + h(C_NFC1))), 256, 0, st()>>>(s_, fast);
  tend();
  apply();
}

// s_loops / s_loops_join (src/ExtreM.cpp:2838-2886) with do_fix, or the bare
// get_wrong_index_max / _min of main.cpp:434,441.
// known: the side's listed-side count when the last counter read holds it
// (no CP pass since), else -1.  With gate, a one-block run also opens
// saddle_loops' gate (returns true then).
bool Run::s_loops(int side, int do_fix, int known, bool gate) {
  bool gated = false;
  int n = side == 2 ? s_.n2 : s_.n1;
  if (do_fix) counts_[EXTREMZ_STAT_S_LOOP_CALLS]++;
  if (!n) {  // (else the first wrong-index launch clears them)
    if (side == 2) set_ctr2(C_WMC, C_WMC2, 0);
    else set_ctr2(C_WMINC, C_WMINC2, 0);
  } else if (known >= 0 && known <= kWrongOne) {  // few listed sides: one block
    tbegin(EXTREMZ_K_PROPOSE);
    k_wrong_index_one<<<1, 1024, 0, st()>>>(s_, side, do_fix,
                                            gate ? (side == 2 ? C_WSC : C_WSCJ) : -1);
    tend();
    gated = gate;
  } else {
    const int b = std::min(blocks_for(n), 512);  // over the listed sides only
    k_wrong_index<<<b, 256, 0, st()>>>(s_, side, 0, 0);
    tbegin(EXTREMZ_K_PROPOSE);
    k_wrong_index<<<b, 256, 0, st()>>>(s_, side, 1, do_fix);
    tend();
  }
  if (do_fix) apply();
  return gated;
}

// saddle_loops / saddle_loops_join (src/ExtreM.cpp:2888-2916) with do_fix,
// or the bare get_wrong_index_saddles(_join) of main.cpp:435,442.
void Run::saddle_loops(int side, int do_fix) {
  int n = side == 2 ? s_.nmax : s_.nmin;
  set_ctr(side == 2 ? C_WSC : C_WSCJ, 0);
  if (do_fix) counts_[EXTREMZ_STAT_SADDLE_LOOP_CALLS]++;
  if (n) {
    tbegin(EXTREMZ_K_PROPOSE);
    k_wrong_saddle<<<blocks_for(n), 256, 0, st()>>>(s_, side, do_fix, -1);
    tend();
  }
  if (do_fix) apply();
}

// saddle_loops(side, 1) when the side's wrong-index count is zero, decided on
// the device (no counter read); the calls made are counted there (C_SLC).
// gated: the gate is open or closed already (by the one-block wrong-index run).
void Run::saddle_loops_if(int side, bool gated) {
  const int n = side == 2 ? s_.nmax : s_.nmin;
  const int gate = side == 2 ? C_WMC : C_WMINC;
  if (!gated) k_saddle_gate<<<1, 64, 0, st()>>>(s_.ctr, gate, side == 2 ? C_WSC : C_WSCJ);
  if (n) {
    tbegin(EXTREMZ_K_PROPOSE);
    k_wrong_saddle<<<blocks_for(n), 256, 0, st()>>>(s_, side, 1, gate);
    tend();
  }
  apply();
}

void Run::classify(const double* d_f, int W, int H, int D, uint8_t* d_t, uint16_t* d_l) {
  setup_geo(W, H, D);
  if (c_->lut.p == nullptr) throw HipError("context has no LUT");
  tbegin(EXTREMZ_K_CLASSIFY);
  k_classify<<<row_blocks(), dim3(64, 4, 1), 0, st()>>>(s_.g, c_->lut.p, d_f, d_t, d_l);
  tend();
  ez::hip_check(hipGetLastError(), "k_classify");
  ez::hip_check(hipStreamSynchronize(st()), "sync");
}

int Run::labels(const double* d_f, int W, int H, int D, int* d_ds, int* d_as, uint8_t* d_t,
                uint16_t* d_l) {
  setup_geo(W, H, D);
  c_->ctr.ensure(C_COUNT_);
  s_.ctr = c_->ctr.p;
  if (d_t && c_->lut.p == nullptr) throw HipError("context has no LUT");
  label_pass(d_f, d_ds, d_as, d_l, nullptr, d_t);
  ez::hip_check(hipGetLastError(), "labels");
  ez::hip_check(hipStreamSynchronize(st()), "sync");
  return 0;
}

// compute_MergeT (src/ExtreM.cpp:2930-2952) as main.cpp:454-463 runs it
// after the correction, on one field with its own types, labels and value
// order: the triplets and the rounds of constructPersistencePairs
// (src/ExtreM.cpp:2132-2249) on the device, constructMergeTree
// (src/ExtreM.cpp:1863-1918) and the branch lists on the host (one entry per
// maximum).  Ties in value are ranked by vertex id; maxima never paired
// keep the reference's zero-initialised (saddle index 0, rank 0) record.
void Run::split_tree(const double* d_f, int W, int H, int D, TreeOut& out) {
  setup_geo(W, H, D);
  const long long N = s_.g.N;
  extremz_ctx* c = c_;
  hipStream_t q = st();
  c->otype.ensure(N);
  c->olow.ensure(N);
  c->oDS.ensure(N);
  c->oAS.ensure(N);
  for (DevBuf<int>* b : {&c->S2, &c->S1, &c->MX, &c->MN, &c->ext_id}) b->ensure(N);
  c->ctr.ensure(C_COUNT_);
  s_ = St{};
  setup_geo(W, H, D);
  s_.lut = c->lut.p;
  s_.otype = c->otype.p;
  s_.olow = c->olow.p;
  s_.S2 = c->S2.p;
  s_.S1 = c->S1.p;
  s_.MX = c->MX.p;
  s_.MN = c->MN.p;
  s_.ext_id = c->ext_id.p;
  s_.ctr = c->ctr.p;
  hip_check(hipMemsetAsync(s_.ctr, 0, sizeof(int) * C_COUNT_, q), "memset");
  k_classify<<<row_blocks(), dim3(64, 4, 1), 0, q>>>(s_.g, s_.lut, d_f, s_.otype, s_.olow);
  label_pass(d_f, c->oDS.p, c->oAS.p, nullptr);
  list_compaction(0);
  argext(d_f, 0);
  pull();
  const int n2 = h(C_N2), nmax = h(C_NMAX), gmin = h(C_GMIN);
  out = TreeOut{};
  if (nmax == 0) return;
  s_.nmax = nmax;
  s_.nmin = 0;
  k_fill_i32<<<blocks_for(N), 256, 0, q>>>(s_.ext_id, -1, N);
  k_ext_index<<<blocks_for(nmax), 256, 0, q>>>(s_);
  DevBuf<long long> mkey, skey, bestkey, bestpos;
  DevBuf<int> T, cnt, pair_pos, parent, ptr;
  mkey.ensure(nmax);
  skey.ensure((size_t)n2 + 1);
  bestkey.ensure(nmax);
  T.ensure((size_t)n2 * kNbr + 1);
  cnt.ensure((size_t)n2 + 1);
  for (DevBuf<int>* b : {&pair_pos, &parent, &ptr}) b->ensure(nmax);
  bestpos.ensure(nmax);
  k_list_keys<<<blocks_for(nmax), 256, 0, q>>>(s_.MX, nmax, d_f, mkey.p);
  if (n2) k_list_keys<<<blocks_for(n2), 256, 0, q>>>(s_.S2, n2, d_f, skey.p);
  std::vector<int> hMX(nmax), hS2(n2);
  std::vector<long long> hmk(nmax), hsk(n2);
  hip_check(hipMemcpyAsync(hMX.data(), s_.MX, 4LL * nmax, hipMemcpyDeviceToHost, q), "d2h");
  hip_check(hipMemcpyAsync(hmk.data(), mkey.p, 8LL * nmax, hipMemcpyDeviceToHost, q), "d2h");
  if (n2) {
    hip_check(hipMemcpyAsync(hS2.data(), s_.S2, 4LL * n2, hipMemcpyDeviceToHost, q), "d2h");
    hip_check(hipMemcpyAsync(hsk.data(), skey.p, 8LL * n2, hipMemcpyDeviceToHost, q), "d2h");
  }
  hip_check(hipStreamSynchronize(q), "sync");
  auto max_less = [&](int a, int b) {
    return hmk[a] < hmk[b] || (hmk[a] == hmk[b] && hMX[a] < hMX[b]);
  };
  int gmax = 0;
  for (int p = 1; p < nmax; p++)
    if (max_less(gmax, p)) gmax = p;
  std::vector<int> iota(nmax);
  for (int p = 0; p < nmax; p++) iota[p] = p;
  hip_check(hipMemcpyAsync(ptr.p, iota.data(), 4LL * nmax, hipMemcpyHostToDevice, q), "h2d");
  k_fill_i32<<<blocks_for(nmax), 256, 0, q>>>(pair_pos.p, -1, nmax);
  k_fill_i32<<<blocks_for(nmax), 256, 0, q>>>(parent.p, -1, nmax);
  if (n2)
    k_tree_triplets<<<blocks_for(n2), 256, 0, q>>>(s_.g, d_f, s_.S2, n2, c->oDS.p, s_.ext_id,
                                                   mkey.p, s_.MX, T.p, cnt.p);
  // constructPersistencePairs: rounds until no maximum is paired
  for (int round = 0;; round++) {
    if (round > nmax + 1) throw HipError("split tree pairing did not finish");
    k_fill_i64<<<blocks_for(nmax), 256, 0, q>>>(bestkey.p, (long long)0x8000000000000000ULL,
                                                nmax);
    k_fill_i64<<<blocks_for(nmax), 256, 0, q>>>(bestpos.p, -1, nmax);
    if (n2) {
      k_tree_best<<<blocks_for(n2), 256, 0, q>>>(T.p, cnt.p, s_.S2, n2, skey.p, gmax, 0,
                                                 bestkey.p, bestpos.p);
      k_tree_best<<<blocks_for(n2), 256, 0, q>>>(T.p, cnt.p, s_.S2, n2, skey.p, gmax, 1,
                                                 bestkey.p, bestpos.p);
    }
    set_ctr(C_CHANGED, 0);
    k_tree_pair<<<blocks_for(nmax), 256, 0, q>>>(T.p, cnt.p, nmax, bestpos.p, pair_pos.p,
                                                 parent.p, ptr.p, s_.ctr + C_CHANGED);
    k_tree_compress<<<blocks_for(nmax), 256, 0, q>>>(nmax, ptr.p);
    if (n2)
      k_tree_relabel<<<blocks_for(n2), 256, 0, q>>>(T.p, cnt.p, n2, ptr.p, mkey.p, s_.MX);
    pull();
    out.rounds++;
    if (!h(C_CHANGED)) break;
  }
  std::vector<int> hpair(nmax), hpar(nmax);
  hip_check(hipMemcpyAsync(hpair.data(), pair_pos.p, 4LL * nmax, hipMemcpyDeviceToHost, q), "d2h");
  hip_check(hipMemcpyAsync(hpar.data(), parent.p, 4LL * nmax, hipMemcpyDeviceToHost, q), "d2h");
  hip_check(hipStreamSynchronize(q), "sync");
  // ranks: ascending (value, id)
  std::vector<int> byrank(nmax), rank(nmax);
  for (int p = 0; p < nmax; p++) byrank[p] = p;
  std::sort(byrank.begin(), byrank.end(), max_less);
  for (int r = 0; r < nmax; r++) rank[byrank[r]] = r;
  // saddle order: descending (value, id) (saddles2 of computeCP); index 0 = largest
  auto sad_larger = [&](int a, int b) {  // list positions in S2
    return hsk[a] > hsk[b] || (hsk[a] == hsk[b] && hS2[a] > hS2[b]);
  };
  int sad0 = -1;
  for (int i = 0; i < n2; i++)
    if (sad0 < 0 || sad_larger(i, sad0)) sad0 = i;
  // maximaTriplets in rank order: {saddle position (-1: global minimum), parent rank}
  std::vector<std::pair<int, int>> mt(nmax);
  out.max_v.resize(nmax);
  out.pair_v.resize(nmax);
  std::vector<char> unpaired(nmax, 0);
  for (int r = 0; r < nmax; r++) {
    const int p = byrank[r];
    out.max_v[r] = hMX[p];
    if (p == gmax) {
      out.pair_v[r] = gmin;
      mt[r] = {-1, r};
    } else if (hpair[p] >= 0) {
      out.pair_v[r] = hS2[hpair[p]];
      mt[r] = {hpair[p], rank[hpar[p]]};
    } else {  // zero-initialised record of the reference
      out.pair_v[r] = 0;
      mt[r] = {sad0, 0};
      unpaired[r] = 1;
    }
  }
  // constructMergeTree: follow larger saddles to the parent
  for (bool same = false; !same;) {
    same = true;
    for (int i = 0; i < nmax; i++) {
      const auto next = mt[mt[i].second];
      if (next.second != mt[i].second && next.first >= 0 && mt[i].first >= 0 &&
          next.first != mt[i].first && sad_larger(next.first, mt[i].first)) {
        mt[i].second = next.second;
        same = false;
      }
    }
  }
  // branches: the maximum, its saddle (or the root marker) and the saddles
  // of its children, descending by (value, id)
  struct Ent {
    long long key;
    int id;
  };
  std::vector<std::vector<Ent>> br(nmax);
  const Ent marker{std::numeric_limits<long long>::min(), 0};
  // a maximum paired with a saddle whose triplet held only itself is its own
  // parent; the reference then lists the saddle's index in saddles2 (sorted
  // by descending (value, id)) as the root marker
  std::vector<int> sad_index;
  auto index_of = [&](int sp) {
    if (sad_index.empty()) {
      std::vector<int> o(n2);
      for (int i = 0; i < n2; i++) o[i] = i;
      std::sort(o.begin(), o.end(), sad_larger);
      sad_index.resize(n2);
      for (int i = 0; i < n2; i++) sad_index[o[i]] = i;
    }
    return sad_index[sp];
  };
  for (int b = 0; b < nmax; b++) {
    br[b].push_back({hmk[byrank[b]], out.max_v[b]});
    const int par = mt[b].second;
    if (par != b && mt[b].first >= 0) {
      const int sp = mt[b].first;
      const Ent e{hsk[sp], hS2[sp]};
      br[b].push_back(e);
      br[par].push_back(e);
    } else {
      Ent m = marker;
      m.id = byrank[b] == gmax ? gmin : (unpaired[b] ? 0 : index_of(mt[b].first));
      br[b].push_back(m);
    }
  }
  out.parent_v.resize(nmax);
  out.branch.resize(nmax);
  for (int b = 0; b < nmax; b++) {
    out.parent_v[b] = out.max_v[mt[b].second];
    auto& v = br[b];
    std::sort(v.begin(), v.end(), [](const Ent& a, const Ent& e) {
      return a.key > e.key || (a.key == e.key && a.id > e.id);
    });
    for (const Ent& e : v) out.branch[b].push_back(e.id);
  }
}

// main.cpp:65-131: checks, buffers, the original field's classification,
// labels, CP lists and largest / smallest saddle maps, and the pre-snap.
int Run::begin(const double* d_orig, double* d_decp, int W, int H, int D, double xi,
               uint8_t* d_counts, bool presnap) {
  if (planar_ && D != 1) {
    g_err = "the plane adjacency (EXTREMZ_ADJ_PLANE4) needs a field of depth 1";
    return EXTREMZ_EINVAL;
  }
  setup_geo(W, H, D);
  const long long N = s_.g.N;
  allocate(N);
  extremz_ctx* c = c_;
  // the float32 fork keeps bound in float (src/ExtreM_cuda_f.cu:124)
  if (f32_) xi = (double)(float)xi;
  s_.xi = xi;
  s_.step_q = step_q_;
  s_.f32 = f32_ ? 1 : 0;
  s_.thr = thr_;
  // with the persistence filter an original extremum -> saddle index lacks
  // the pairs the filter dropped, through which a decp change could reach a
  // side: every CP pass of decp is a full pass
  if (thr_ > 0.0) inc_off_ = true;
  s_.step = step_q_ ? -xi / step_q_ : (f32_ ? (double)(-(float)xi / 4.0f) : 0.0);
  {
    double sd = -4.0 * xi;
    long long b;
    std::memcpy(&b, &sd, 8);
    s_.sent = b >= 0 ? b : (b ^ 0x7fffffffffffffffLL);
  }
  s_.lut = c->lut.p;
  s_.orig = d_orig;
  s_.decp = d_decp;
  s_.delta = c->delta.p;
  s_.cnt = d_counts ? d_counts : c->cnt.p;
  s_.otype = c->otype.p;
  s_.olow = c->olow.p;
  s_.dlow = c->dlow.p;
  s_.oDS = c->oDS.p;
  s_.oAS = c->oAS.p;
  s_.dDS = c->dDS.p;
  s_.dAS = c->dAS.p;
  s_.stamp = c->stamp.p;
  s_.dflag = c->dflag.p;
  s_.touched = c->touched.p;
  s_.dirty = c->dirty.p;
  s_.ctr = c->ctr.p;
  s_.tctr = C_TOUCHED;
  s_.dctr = C_DIRTY;
  s_.S2 = c->S2.p;
  s_.S1 = c->S1.p;
  s_.MX = c->MX.p;
  s_.MN = c->MN.p;
  s_.of2 = c->of2.p;
  s_.ol2 = c->ol2.p;
  s_.of1 = c->of1.p;
  s_.ol1 = c->ol1.p;
  s_.df2 = c->df2.p;
  s_.dl2 = c->dl2.p;
  s_.df1 = c->df1.p;
  s_.dl1 = c->dl1.p;
  s_.olarg = c->olarg.p;
  s_.osmall = c->osmall.p;
  s_.dlarg = c->dlarg.p;
  s_.dsmall = c->dsmall.p;
  s_.act0 = c->act0.p;
  s_.ext_id = c->ext_id.p;
  s_.wn2 = c->wn2.p;
  s_.wn1 = c->wn1.p;
  s_.wnl2 = c->wnl2.p;
  s_.wnl1 = c->wnl1.p;
  s_.minsad2 = c->minsad2.p;
  s_.minsad1 = c->minsad1.p;
  s_.pdir = c->pdir.p;
  s_.opdir = c->opdir.p;
  s_.lflag = c->lflag.p;
  s_.lchg = c->lchg.p;
  s_.cflag = c->cflag.p;
  s_.tu = c->tu.p;
  s_.td = c->td.p;
  s_.spos2 = c->spos2.p;
  s_.spos1 = c->spos1.p;
  s_.qu = c->qu.p;
  s_.qd = c->qd.p;
  {
    const long long cap = inc_cap(N);
    s_.icap = (int)cap;
    int* il = c->ilist.p;
    for (int** l : {&s_.al, &s_.cu, &s_.cd, &s_.ou, &s_.od}) {
      *l = il;
      il += cap;
    }
  }

  hipStream_t q = st();
  const int bN = blocks_for(N);
  ez::hip_check(hipMemsetAsync(s_.ctr, 0, sizeof(int) * C_COUNT_, q), "memset");
  // main.cpp:67-75 (the global extrema of orig), with the check of orig
  argext(d_orig, 1);
  k_check_finite<<<bN, 256, 0, q>>>(d_decp, N, s_.ctr, s_.f32);
  k_fill_i64<<<bN, 256, 0, q>>>(s_.delta, s_.sent, N);
  ez::hip_check(hipMemsetAsync(s_.cnt, 0, N, q), "memset");
  ez::hip_check(hipMemsetAsync(s_.stamp, 0, sizeof(int) * N, q), "memset");
  ez::hip_check(hipMemsetAsync(s_.dflag, 0, N, q), "memset");
  ez::hip_check(hipMemsetAsync(s_.wn2, 0, N, q), "memset");
  ez::hip_check(hipMemsetAsync(s_.wn1, 0, N, q), "memset");
  for (uint8_t* p : {s_.lflag, s_.cflag, s_.tu, s_.td})
    ez::hip_check(hipMemsetAsync(p, 0, N, q), "memset");
  for (int* p : {s_.minsad2, s_.minsad1})
    k_fill_i32<<<bN, 256, 0, q>>>(p, 0x7fffffff, N);
  pull();
  if (h(C_FAIL)) return EXTREMZ_EINVAL;

  // main.cpp:84-88: the original field's classification (lower masks and
  // types) comes out of its own label pass, which also makes the pre-snap
  // (main.cpp:74-75, 122-131; a caller that drives the steps snaps itself)
  label_pass(d_orig, s_.oDS, s_.oAS, s_.olow, s_.opdir, s_.otype,
             presnap ? Presnap{s_.decp, s_.xi, s_.f32, s_.ctr} : Presnap{});
  // main.cpp:100-119: CP lists, triplets, largest/smallest saddle maps of orig
  // with the saddle positions and the extremum index
  list_compaction(1);
  pull();
  s_.n2 = h(C_N2);
  s_.n1 = h(C_N1);
  s_.nmax = h(C_NMAX);
  s_.nmin = h(C_NMIN);
  counts_[EXTREMZ_STAT_N_SADDLE2] = s_.n2;
  counts_[EXTREMZ_STAT_N_SADDLE1] = s_.n1;
  counts_[EXTREMZ_STAT_N_MAX] = s_.nmax;
  counts_[EXTREMZ_STAT_N_MIN] = s_.nmin;
  // every saddle records at most one pair per distinct root of its star, in
  // the region of its 256-side chunk
  const long long nchunk = ((long long)s_.n2 + s_.n1) / 256 + 4;
  c->pairs.ensure(nchunk * kChunkPairs);
  c->pcnt.ensure(nchunk);
  {  // incremental CP passes: saddle positions, side and slot lists and stamps
    const long long n21 = (long long)s_.n2 + s_.n1, ns = (long long)s_.nmax + s_.nmin;
    c->sstamp.ensure(n21 + 1);
    c->slists.ensure(4 * n21 + 1);
    c->wflag.ensure(n21 + 1);
    s_.wflag = c->wflag.p;
    c->slstamp.ensure(ns + 1);
    c->dslot.ensure(ns + 1);
    c->pstamp.ensure(ns + 1);
    c->plists.ensure(2 * (ns + 1));
    s_.pstamp = c->pstamp.p;
    s_.pprev = c->plists.p;
    s_.pnext = c->plists.p + ns + 1;
    s_.sstamp = c->sstamp.p;
    s_.slstamp = c->slstamp.p;
    s_.dslot = c->dslot.p;
    int* sl = c->slists.p;
    s_.dsl2 = sl;
    s_.dsl1 = sl + s_.n2;
    s_.cl2 = sl + n21;
    s_.cl1 = s_.cl2 + s_.n2;
    s_.cln2 = sl + 2 * n21;
    s_.cln1 = s_.cln2 + s_.n2;
    s_.wl2 = sl + 3 * n21;
    s_.wl1 = s_.wl2 + s_.n2;
    ez::hip_check(hipMemsetAsync(s_.sstamp, 0, sizeof(int) * (n21 + 1), q), "memset");
    ez::hip_check(hipMemsetAsync(s_.slstamp, 0, sizeof(int) * (ns + 1), q), "memset");
    ez::hip_check(hipMemsetAsync(s_.pstamp, 0, sizeof(int) * (ns + 1), q), "memset");
    // first-saddle-wins words of the wrong-index steps, one per slot
    c->win_a.ensure(ns + 1);
    c->win_b.ensure(ns + 1);
    s_.win_a = c->win_a.p;
    s_.win_b = c->win_b.p;
    k_fill_i32<<<blocks_for(ns + 1), 256, 0, q>>>(s_.win_a, 0x7fffffff, ns + 1);
    k_fill_i32<<<blocks_for(ns + 1), 256, 0, q>>>(s_.win_b, 0x7fffffff, ns + 1);
  }
  {
    const long long ns = (long long)s_.nmax + s_.nmin;
    c->csr_off.ensure(ns + 1);
    c->csr_cur.ensure(ns + 1);
    c->pid.ensure(ns + 1);
    c->pkey.ensure(ns + 1);
    c->longs.ensure(ns + 1);
    s_.longs = c->longs.p;
    c->pv2.ensure((size_t)s_.n2 + 1);
    c->pv1.ensure((size_t)s_.n1 + 1);
    s_.pv2 = c->pv2.p;
    s_.pv1 = c->pv1.p;
    s_.csr_off = c->csr_off.p;
    s_.csr_cur = c->csr_cur.p;
  }

  slot_pass(0, d_orig, s_.oDS, s_.oAS, s_.olow, s_.of2, s_.ol2, s_.of1, s_.ol1, s_.olarg,
            s_.osmall);
  return 0;
}

int Run::correct(const double* d_orig, double* d_decp, int W, int H, int D, double xi,
                 uint8_t* d_counts) {
  if (int rc = begin(d_orig, d_decp, W, H, D, xi, d_counts)) return rc;
  const long long N = s_.g.N;
  hipStream_t q = st();
  const int bN = blocks_for(N);

  c_loops();     // main.cpp:139
  cp_dec(true);  // main.cpp:172-191
  wrong_neighbors();  // main.cpp:210-226
  int nfc = h(C_NFC), nfc1 = h(C_NFC1);
  // (wrong_neighbors and c_loops end with a counter read: cp_dec can use it)
  if (nfc == 0) {  // main.cpp:230-254
    cp_dec(false, true);
    s_loops(2, 1);
    saddle_loops_if(2);
  }
  if (nfc1 == 0) {  // main.cpp:255-274
    cp_dec(false, nfc != 0);
    s_loops(1, 1);
    saddle_loops_if(1);
  }
  auto any = [&]() {  // on the last counter read
    return h(C_NFC) > 0 || h(C_NFC1) > 0 || h(C_WMC) > 0 || h(C_WMC2) > 0 || h(C_WSC) > 0 ||
           h(C_WMINC) > 0 || h(C_WMINC2) > 0 || h(C_WSCJ) > 0;
  };
  pull();
  // entered: the next iteration's r_loops and first c_loops entry have run,
  // and the last read came after them
  bool entered = false;
  while (any()) {  // main.cpp:277-446
    if (++counts_[EXTREMZ_STAT_OUTER_ITERS] > max_outer_) throw std::range_error("outer");
    if (!entered) r_loops();
    for (int rep = 0; rep < 3; rep++) {
      c_loops(entered && rep == 0);
      cp_dec(false, true);
      wrong_neighbors();
      nfc = h(C_NFC);
      nfc1 = h(C_NFC1);
      // (the listed-side counts of that read stand: no CP pass follows it here)
      const int wl2 = h(C_WL2), wl1 = h(C_WL1);
      if (rep < 2) {
        if (nfc == 0) {
          saddle_loops_if(2, s_loops(2, 1, wl2, true));
        }
        if (nfc1 == 0) {
          saddle_loops_if(1, s_loops(1, 1, wl1, true));
        }
      } else {
        if (nfc == 0) {
          s_loops(2, 0, wl2);
          saddle_loops(2, 0);
        }
        if (nfc1 == 0) {
          s_loops(1, 0, wl1);
          saddle_loops(1, 0);
        }
      }
    }
    // The next iteration's r_loops and c_loops entry go before the loop test.
    // When the loop ends every branch counter is zero, and no edit has
    // happened since the last c_loops: fixpath has no wrong neighbour and the
    // dirty list is empty, so both do nothing.  One read then serves the test
    // and the first frontier length.
    r_loops();
    c_loops_entry();
    pull();
    entered = true;
  }
  if (entered) {  // the last such start began no iteration
    counts_[EXTREMZ_STAT_R_LOOP_CALLS]--;
    counts_[EXTREMZ_STAT_C_LOOPS_CALLS]--;
  }
  set_ctr(C_EDITED, 0);
  k_count_edited<<<bN, 256, 0, q>>>(s_.cnt, N, s_.ctr);
  pull();
  counts_[EXTREMZ_STAT_EDITED] = h(C_EDITED);
  counts_[EXTREMZ_STAT_SADDLE_LOOP_CALLS] += h(C_SLC);
  return 0;
}

static std::mutex g_cache_mu;
static std::vector<extremz_ctx*> g_cache;

extremz_ctx* new_ctx(int device) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return nullptr;
  }
  ez::hip_check(hipSetDevice(device), "hipSetDevice");
  hipDeviceProp_t prop;
  ez::hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return nullptr;
  extremz_ctx* c = new extremz_ctx();
  c->device = device;
  ez::hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  ez::hip_check(hipHostMalloc((void**)&c->h_ctr, sizeof(int) * (C_COUNT_ + 1),
                             hipHostMallocCoherent),
                "hipHostMalloc");
  std::memset(c->h_ctr, 0, sizeof(int) * (C_COUNT_ + 1));
  const LutTable& L = lut_table();
  c->lut.ensure(L.table.size());
  ez::hip_check(hipMemcpy(c->lut.p, L.table.data(), L.table.size(), hipMemcpyHostToDevice), "lut");
  {
    static std::mutex mu;  // contexts may be created from several threads
    std::lock_guard<std::mutex> lk(mu);
    upload_constants(kDx, kDy, kDz, L.cls_key, L.cls_off);
  }
  return c;
}

void free_ctx(extremz_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& p : c->ev_pool) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  c->each_buf([](auto& b) { b.release(); });
  c->lut.release();
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

size_t ctx_bytes(extremz_ctx* c) {
  size_t n = 0;
  c->each_buf([&](auto& b) { n += b.bytes(); });
  return n;
}

void ctx_release(extremz_ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  c->each_buf([](auto& b) { b.release(); });
}

static void fill_stats(extremz_ctx* c, const Run& r, int status, double wall_ms,
                       extremz_stats* stats) {
  if (!stats) return;
  std::memset(stats, 0, sizeof(*stats));
  for (int i = 0; i < EXTREMZ_STAT_COUNT_; i++) stats->counts[i] = r.counts_[i];
  stats->counts[EXTREMZ_STAT_STATUS] = status;
  stats->counts[EXTREMZ_STAT_HOST_SYNCS] = c->host_syncs;
  stats->wall_ms = wall_ms;
  if (c->timing && c->ev_used) {
    (void)hipStreamSynchronize(c->stream);
    for (size_t i = 0; i < c->ev_used; i++) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c->ev_pool[i].a, c->ev_pool[i].b) == hipSuccess) {
        stats->kernel_ms[c->ev_pool[i].fam] += ms;
        stats->kernel_launches[c->ev_pool[i].fam]++;
      }
    }
  }
}

bool valid_opts(const extremz_opts* o) {
  return !o || (o->fixed_step_q >= 0 && o->fixed_step_q <= 254 &&  // counts are uint8
                (o->adjacency == EXTREMZ_ADJ_STENCIL14 || o->adjacency == EXTREMZ_ADJ_PLANE4) &&
                (o->rule_set == EXTREMZ_RULES_CPU ||
                 (o->rule_set == EXTREMZ_RULES_F32 && o->fixed_step_q == 0)) &&
                std::isfinite(o->persistence) && o->persistence >= 0.0);
}

bool valid_shape(int W, int H, int D) {
  if (W <= 0 || H <= 0 || D <= 0) return false;
  long long N = (long long)W * H * D;
  return N <= (1LL << 30);  // entries 2*v + side must fit an int
}

std::mutex& cache_mutex() { return g_cache_mu; }
int ctx_device(extremz_ctx* c) { return c->device; }
hipStream_t ctx_stream(extremz_ctx* c) { return c->stream; }

extremz_ctx* cached_ctx(int device) {
  for (extremz_ctx* c : g_cache)
    if (c->device == device) return c;
  extremz_ctx* c = new_ctx(device);
  if (c) g_cache.push_back(c);
  return c;
}

int run_correct(extremz_ctx* c, const double* d_orig, double* d_decp, int W, int H, int D,
                double xi, const extremz_opts* opts, extremz_stats* stats, uint8_t* d_counts) {
  auto t0 = std::chrono::steady_clock::now();
  ez::hip_check(hipSetDevice(c->device), "hipSetDevice");
  c->host_syncs = 0;
  Run r(c, opts);
  int status = guarded([&] { return r.correct(d_orig, d_decp, W, H, D, xi, d_counts); });
  double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  fill_stats(c, r, status, ms, stats);
  return status;
}

static extremz_opts adjacency_opts(int planar) {
  extremz_opts o{};
  o.adjacency = planar ? EXTREMZ_ADJ_PLANE4 : EXTREMZ_ADJ_STENCIL14;
  return o;
}

void field_classify(extremz_ctx* c, int planar, const double* d_f, int W, int H, int D,
                    uint8_t* d_types, uint16_t* d_lower) {
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  const extremz_opts o = adjacency_opts(planar);
  Run r(c, &o);
  r.classify(d_f, W, H, D, d_types, d_lower);
}

void field_labels(extremz_ctx* c, int planar, const double* d_f, int W, int H, int D, int* d_ds,
                  int* d_as) {
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  const extremz_opts o = adjacency_opts(planar);
  Run r(c, &o);
  r.labels(d_f, W, H, D, d_ds, d_as);
}

void field_split_tree(extremz_ctx* c, int planar, const double* d_f, int W, int H, int D,
                      TreeOut& out) {
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  const extremz_opts o = adjacency_opts(planar);
  Run r(c, &o);
  r.split_tree(d_f, W, H, D, out);
}

// The correction's steps for a caller that drives them (friend of Run).
struct Stepper {
  extremz_opts opts;
  Run run;
  bool begun = false;
  Stepper(extremz_ctx* c, const extremz_opts& o) : opts(o), run(c, &opts) {}

  Run& r() {
    if (!begun) throw HipError("correction step before its set-up");
    return run;
  }
  void begin(const double* d_orig, double* d_decp, uint8_t* d_counts, int W, int H, int D,
             double xi) {
    hip_check(hipSetDevice(run.c_->device), "hipSetDevice");
    run.c_->host_syncs = 0;
    const int rc = run.begin(d_orig, d_decp, W, H, D, xi, d_counts, false);
    if (rc) throw HipError(g_err.empty() ? "correction set-up failed" : g_err);
    begun = true;
  }
  void c_loops() { r().c_loops(false); }
  void cp() { r().cp_dec(!run.inc_ok_, false); }
  void candidates(std::vector<int2>& out) {
    Run& x = r();
    x.pull();
    const int n = x.h(C_ACT0);
    out.resize(n);
    if (n)
      hip_check(hipMemcpyAsync(out.data(), x.c_->act0.p, sizeof(int2) * n,
                               hipMemcpyDeviceToHost, x.st()),
                "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(x.st()), "hipStreamSynchronize");
  }
  void r_loops(const int* l2, int n2, const int* l1, int n1) {
    Run& x = r();
    if (n2 < 0 || n1 < 0 || n2 > x.s_.g.N || n1 > x.s_.g.N)
      throw HipError("wrong-neighbour list longer than the field");
    hipStream_t q = x.st();
    if (n2) hip_check(hipMemcpyAsync(x.s_.wnl2, l2, 4LL * n2, hipMemcpyHostToDevice, q), "h2d");
    if (n1) hip_check(hipMemcpyAsync(x.s_.wnl1, l1, 4LL * n1, hipMemcpyHostToDevice, q), "h2d");
    // the lists are the caller's: the device's own selection is not used
    x.set_ctr(C_NFC, n2);
    x.set_ctr(C_NFC1, n1);
    x.pull();  // whether decp changed since its last CP pass (fixpath's fast path)
    x.r_loops();
    x.pull();
  }
  void s_loops(int side, int do_fix) {
    r().s_loops(side, do_fix);
    run.pull();
  }
  void saddle_loops(int side, int do_fix) {
    r().saddle_loops(side, do_fix);
    run.pull();
  }
  void counters(int out[6]) {
    Run& x = r();
    x.pull();
    const int idx[6] = {C_WMC, C_WMC2, C_WMINC, C_WMINC2, C_WSC, C_WSCJ};
    for (int k = 0; k < 6; k++) out[k] = x.h(idx[k]);
  }
  void decp_replaced() {
    Run& x = r();
    x.inc_ok_ = false;
    x.clean_ = false;
    x.set_ctr(C_ANY_CHANGE, 1);
  }
  const uint8_t* types() { return r().s_.otype; }
  void labels(const int** ds, const int** as) {
    *ds = r().s_.dDS;
    *as = run.s_.dAS;
  }
};

Stepper* stepper_new(extremz_ctx* c, int planar, int max_c_iters) {
  extremz_opts o = adjacency_opts(planar);
  o.max_c_iters = max_c_iters;
  return new Stepper(c, o);
}
void stepper_free(Stepper* s) { delete s; }
void stepper_begin(Stepper* s, const double* d_orig, double* d_decp, uint8_t* d_counts, int W,
                   int H, int D, double xi) {
  s->begin(d_orig, d_decp, d_counts, W, H, D, xi);
}
void stepper_c_loops(Stepper* s) { s->c_loops(); }
void stepper_cp(Stepper* s) { s->cp(); }
void stepper_candidates(Stepper* s, std::vector<int2>& out) { s->candidates(out); }
void stepper_r_loops(Stepper* s, const int* l2, int n2, const int* l1, int n1) {
  s->r_loops(l2, n2, l1, n1);
}
void stepper_s_loops(Stepper* s, int side, int do_fix) { s->s_loops(side, do_fix); }
void stepper_saddle_loops(Stepper* s, int side, int do_fix) { s->saddle_loops(side, do_fix); }
void stepper_counters(Stepper* s, int out[6]) { s->counters(out); }
void stepper_decp_replaced(Stepper* s) { s->decp_replaced(); }
const uint8_t* stepper_types(Stepper* s) { return s->types(); }
void stepper_labels(Stepper* s, const int** d_ds, const int** d_as) { s->labels(d_ds, d_as); }

int correct_staged(extremz_ctx* c, const void* orig, int orig_dtype, bool on_host, double* decp,
                   uint8_t* counts, int W, int H, int D, double xi, const extremz_opts* opts,
                   extremz_stats* stats) {
  hip_check(hipSetDevice(c->device), "hipSetDevice");
  const size_t N = (size_t)W * H * D;
  hipStream_t q = c->stream;
  const double* d_orig = (const double*)orig;
  double* d_decp = decp;
  uint8_t* d_counts = counts;
  if (on_host) {
    c->decp_tmp.ensure(N);
    c->counts_tmp.ensure(N);
    d_decp = c->decp_tmp.p;
    d_counts = c->counts_tmp.p;
  }
  if (on_host || orig_dtype == EXTREMZ_DTYPE_F32) {
    c->orig_tmp.ensure(N);
    if (!on_host) {
      k_f32_to_f64<<<2048, 256, 0, q>>>((const float*)orig, c->orig_tmp.p, (long long)N);
    } else if (orig_dtype == EXTREMZ_DTYPE_F64) {
      hip_check(hipMemcpyAsync(c->orig_tmp.p, orig, N * 8, hipMemcpyHostToDevice, q), "h2d");
    } else {
      // stage float32 in the decp buffer, then promote exactly
      hip_check(hipMemcpyAsync(c->decp_tmp.p, orig, N * 4, hipMemcpyHostToDevice, q), "h2d");
      k_f32_to_f64<<<2048, 256, 0, q>>>((const float*)c->decp_tmp.p, c->orig_tmp.p,
                                        (long long)N);
    }
    d_orig = c->orig_tmp.p;
  }
  if (on_host)
    hip_check(hipMemcpyAsync(d_decp, decp, N * 8, hipMemcpyHostToDevice, q), "h2d");
  int rc = run_correct(c, d_orig, d_decp, W, H, D, xi, opts, stats, d_counts);
  if (on_host) {
    hip_check(hipMemcpyAsync(decp, d_decp, N * 8, hipMemcpyDeviceToHost, q), "d2h");
    if (counts) hip_check(hipMemcpyAsync(counts, d_counts, N, hipMemcpyDeviceToHost, q), "d2h");
  }
  hip_check(hipStreamSynchronize(q), "sync");
  return rc;
}

}  // namespace ez

// ================================================================ C ABI ====
extern "C" {

const char* extremz_last_error(void) { return ez::g_err.c_str(); }

int extremz_lut(int32_t* out, int cap) {
  const ez::LutTable& L = ez::lut_table();
  if (cap < L.ref_entries || out == nullptr) return L.ref_entries;
  for (int i = 0; i < L.ref_entries; i++) out[i] = L.table[i];
  return L.ref_entries;
}

int extremz_ctx_create(int device, extremz_ctx** out) {
  if (!out) return EXTREMZ_EINVAL;
  *out = nullptr;
  return ez::guarded([&] {
    extremz_ctx* c = ez::new_ctx(device);
    if (!c) {
      ez::g_err = "no gfx950 device with that ordinal";
      return EXTREMZ_ENODEV;
    }
    *out = c;
    return EXTREMZ_OK;
  });
}

int extremz_ctx_destroy(extremz_ctx* ctx) {
  ez::free_ctx(ctx);
  return EXTREMZ_OK;
}

int extremz_correct_device(extremz_ctx* ctx, const double* d_orig, double* d_decp, int W, int H,
                           int D, double xi, const extremz_opts* opts, extremz_stats* stats,
                           uint8_t* d_counts) {
  if (!ctx || !d_orig || !d_decp || !ez::valid_shape(W, H, D) || !(xi > 0) || !std::isfinite(xi) ||
      !ez::valid_opts(opts))
    return EXTREMZ_EINVAL;
  return ez::run_correct(ctx, d_orig, d_decp, W, H, D, xi, opts, stats, d_counts);
}

int extremz_correct(const void* orig, int orig_dtype, double* decp, int W, int H, int D, double xi,
                    const extremz_opts* opts, extremz_stats* stats, uint8_t* counts) {
  if (!orig || !decp || !ez::valid_shape(W, H, D) || !(xi > 0) || !std::isfinite(xi) ||
      (orig_dtype != EXTREMZ_DTYPE_F32 && orig_dtype != EXTREMZ_DTYPE_F64) ||
      !ez::valid_opts(opts))
    return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  return ez::guarded([&] {
    extremz_ctx* c = ez::cached_ctx(0);
    if (!c) {
      ez::g_err = "no gfx950 device";
      return EXTREMZ_ENODEV;
    }
    return ez::correct_staged(c, orig, orig_dtype, true, decp, counts, W, H, D, xi, opts, stats);
  });
}

int extremz_classify(const double* field, int W, int H, int D, uint8_t* types,
                     uint16_t* lower_masks) {
  if (!field || !types || !lower_masks || !ez::valid_shape(W, H, D)) return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  return ez::guarded([&] {
    extremz_ctx* c = ez::cached_ctx(0);
    if (!c) return EXTREMZ_ENODEV;
    ez::hip_check(hipSetDevice(c->device), "hipSetDevice");
    const size_t N = (size_t)W * H * D;
    c->orig_tmp.ensure(N);
    c->otype.ensure(N);
    c->olow.ensure(N);
    ez::hip_check(hipMemcpyAsync(c->orig_tmp.p, field, N * 8, hipMemcpyHostToDevice, c->stream), "h2d");
    ez::Run r(c, nullptr);
    r.classify(c->orig_tmp.p, W, H, D, c->otype.p, c->olow.p);
    ez::hip_check(hipMemcpyAsync(types, c->otype.p, N, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipMemcpyAsync(lower_masks, c->olow.p, N * 2, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipStreamSynchronize(c->stream), "sync");
    return EXTREMZ_OK;
  });
}

int extremz_labels(const double* field, int W, int H, int D, int32_t* ds, int32_t* as) {
  if (!field || !ds || !as || !ez::valid_shape(W, H, D)) return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  return ez::guarded([&] {
    extremz_ctx* c = ez::cached_ctx(0);
    if (!c) return EXTREMZ_ENODEV;
    ez::hip_check(hipSetDevice(c->device), "hipSetDevice");
    const size_t N = (size_t)W * H * D;
    c->orig_tmp.ensure(N);
    c->oDS.ensure(N);
    c->oAS.ensure(N);
    ez::hip_check(hipMemcpyAsync(c->orig_tmp.p, field, N * 8, hipMemcpyHostToDevice, c->stream), "h2d");
    ez::Run r(c, nullptr);
    r.labels(c->orig_tmp.p, W, H, D, c->oDS.p, c->oAS.p);
    ez::hip_check(hipMemcpyAsync(ds, c->oDS.p, N * 4, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipMemcpyAsync(as, c->oAS.p, N * 4, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipStreamSynchronize(c->stream), "sync");
    return EXTREMZ_OK;
  });
}

int extremz_label_pass(const double* field, int W, int H, int D, int32_t* ds, int32_t* as,
                       uint8_t* types, uint16_t* lower_masks) {
  if (!field || !ds || !as || !types || !lower_masks || !ez::valid_shape(W, H, D))
    return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  return ez::guarded([&] {
    extremz_ctx* c = ez::cached_ctx(0);
    if (!c) return EXTREMZ_ENODEV;
    ez::hip_check(hipSetDevice(c->device), "hipSetDevice");
    const size_t N = (size_t)W * H * D;
    c->orig_tmp.ensure(N);
    c->oDS.ensure(N);
    c->oAS.ensure(N);
    c->otype.ensure(N);
    c->olow.ensure(N);
    ez::hip_check(hipMemcpyAsync(c->orig_tmp.p, field, N * 8, hipMemcpyHostToDevice, c->stream), "h2d");
    ez::Run r(c, nullptr);
    r.labels(c->orig_tmp.p, W, H, D, c->oDS.p, c->oAS.p, c->otype.p, c->olow.p);
    ez::hip_check(hipMemcpyAsync(ds, c->oDS.p, N * 4, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipMemcpyAsync(as, c->oAS.p, N * 4, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipMemcpyAsync(types, c->otype.p, N, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipMemcpyAsync(lower_masks, c->olow.p, N * 2, hipMemcpyDeviceToHost, c->stream), "d2h");
    ez::hip_check(hipStreamSynchronize(c->stream), "sync");
    return EXTREMZ_OK;
  });
}

int extremz_encode(const double* orig, const double* decp_before, const double* decp_after,
                   const uint8_t* counts, int64_t n, double xi, uint8_t* index_bitmap,
                   double* edits, uint64_t* mantissas, int32_t* exponents, int64_t* n_edits) {
  if (!orig || !decp_before || !decp_after || !counts || n <= 0 || n >= (1LL << 31) - 64 ||
      !index_bitmap)
    return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  return ez::guarded([&] {
    extremz_ctx* c = ez::cached_ctx(0);
    if (!c) return EXTREMZ_ENODEV;
    ez::hip_check(hipSetDevice(c->device), "hipSetDevice");
    hipStream_t q = c->stream;
    const long long N = n;
    const long long nbytes = (3 * N + 7) / 8;
    const long long nblk = (N + 1023) / 1024;
    ez::DevBuf<double> o, b, a, e;
    ez::DevBuf<uint8_t> cn, bm;
    ez::DevBuf<int> flag, pos, bs, ex;
    ez::DevBuf<unsigned long long> mt;
    o.ensure(N); b.ensure(N); a.ensure(N); cn.ensure(N); bm.ensure(nbytes);
    flag.ensure(N); pos.ensure(N); bs.ensure(nblk);
    ez::hip_check(hipMemcpyAsync(o.p, orig, N * 8, hipMemcpyHostToDevice, q), "h2d");
    ez::hip_check(hipMemcpyAsync(b.p, decp_before, N * 8, hipMemcpyHostToDevice, q), "h2d");
    ez::hip_check(hipMemcpyAsync(a.p, decp_after, N * 8, hipMemcpyHostToDevice, q), "h2d");
    ez::hip_check(hipMemcpyAsync(cn.p, counts, N, hipMemcpyHostToDevice, q), "h2d");
    ez::k_bitmap3<<<2048, 256, 0, q>>>(b.p, a.p, cn.p, N, bm.p, nbytes);
    ez::k_edit_flags<<<2048, 256, 0, q>>>(b.p, a.p, cn.p, N, flag.p);
    ez::k_scan_block<<<(unsigned)nblk, 1024, 0, q>>>(flag.p, pos.p, N, bs.p);
    std::vector<int> hb(nblk);
    ez::hip_check(hipMemcpyAsync(hb.data(), bs.p, nblk * 4, hipMemcpyDeviceToHost, q), "d2h");
    ez::hip_check(hipMemcpyAsync(index_bitmap, bm.p, nbytes, hipMemcpyDeviceToHost, q), "d2h");
    ez::hip_check(hipStreamSynchronize(q), "sync");
    long long total = 0;
    for (long long i = 0; i < nblk; i++) {
      long long v = hb[i];
      hb[i] = (int)total;
      total += v;
    }
    if (n_edits) *n_edits = total;
    if (total > 0 && (edits || mantissas || exponents)) {
      ez::hip_check(hipMemcpyAsync(bs.p, hb.data(), nblk * 4, hipMemcpyHostToDevice, q), "h2d");
      ez::k_scan_add<<<(unsigned)nblk, 1024, 0, q>>>(pos.p, N, bs.p);
      e.ensure(total);
      mt.ensure(total);
      ex.ensure(total);
      ez::k_edit_write<<<2048, 256, 0, q>>>(o.p, b.p, pos.p, flag.p, N, xi, e.p, mt.p, ex.p);
      if (edits) ez::hip_check(hipMemcpyAsync(edits, e.p, total * 8, hipMemcpyDeviceToHost, q), "d2h");
      if (mantissas)
        ez::hip_check(hipMemcpyAsync(mantissas, mt.p, total * 8, hipMemcpyDeviceToHost, q), "d2h");
      if (exponents)
        ez::hip_check(hipMemcpyAsync(exponents, ex.p, total * 4, hipMemcpyDeviceToHost, q), "d2h");
    }
    ez::hip_check(hipStreamSynchronize(q), "sync");
    return EXTREMZ_OK;
  });
}

static int tree_of(const double* field, int W, int H, int D, ez::TreeOut& t) {
  extremz_ctx* c = ez::cached_ctx(0);
  if (!c) return EXTREMZ_ENODEV;
  ez::hip_check(hipSetDevice(c->device), "hipSetDevice");
  const size_t N = (size_t)W * H * D;
  c->orig_tmp.ensure(N);
  ez::hip_check(hipMemcpyAsync(c->orig_tmp.p, field, N * 8, hipMemcpyHostToDevice, c->stream),
                "h2d");
  ez::Run r(c, nullptr);
  r.split_tree(c->orig_tmp.p, W, H, D, t);
  return EXTREMZ_OK;
}

int extremz_split_tree(const double* field, int W, int H, int D, int32_t* max_v, int32_t* pair_v,
                       int32_t* parent_v, int32_t* branch_off, int32_t* branch_v, int cap_max,
                       int cap_branch, int* rounds) {
  if (!field || !ez::valid_shape(W, H, D)) return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  ez::TreeOut t;
  int rc = ez::guarded([&] { return tree_of(field, W, H, D, t); });
  if (rc != EXTREMZ_OK) return rc;
  const int n = (int)t.max_v.size();
  int total = 0;
  for (auto& b : t.branch) total += (int)b.size();
  if (n > cap_max || total > cap_branch || !max_v || !pair_v || !parent_v || !branch_off ||
      !branch_v)
    return -std::max(n, total) - 1;
  int k = 0;
  for (int i = 0; i < n; i++) {
    max_v[i] = t.max_v[i];
    pair_v[i] = t.pair_v[i];
    parent_v[i] = t.parent_v[i];
    branch_off[i] = k;
    for (int v : t.branch[i]) branch_v[k++] = v;
  }
  branch_off[n] = k;
  if (rounds) *rounds = t.rounds;
  return n;
}

int extremz_compare_split_trees(const double* a, const double* b, int W, int H, int D,
                                extremz_tree_cmp* out) {
  if (!a || !b || !out || !ez::valid_shape(W, H, D)) return EXTREMZ_EINVAL;
  std::lock_guard<std::mutex> lk(ez::g_cache_mu);
  ez::TreeOut ta, tb;
  int rc = ez::guarded([&] {
    int r = tree_of(a, W, H, D, ta);
    return r != EXTREMZ_OK ? r : tree_of(b, W, H, D, tb);
  });
  if (rc != EXTREMZ_OK) return rc;
  const size_t na = ta.max_v.size(), nb = tb.max_v.size(), n = std::min(na, nb);
  *out = extremz_tree_cmp{};
  out->n_max_a = (int64_t)na;
  out->n_max_b = (int64_t)nb;
  out->rounds_a = ta.rounds;
  out->rounds_b = tb.rounds;
  const int64_t extra = (int64_t)(std::max(na, nb) - n);
  out->pair_mismatch = extra;
  out->branch_mismatch = extra;
  for (size_t i = 0; i < n; i++) {
    out->pair_mismatch += ta.pair_v[i] != tb.pair_v[i] || ta.max_v[i] != tb.max_v[i];
    out->branch_mismatch += ta.branch[i] != tb.branch[i];
  }
  return EXTREMZ_OK;
}

}  // extern "C"
