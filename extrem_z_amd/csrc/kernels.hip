// kernels.hip -- gfx950 kernels of the ExtreM correction engine.
//
// One vertex per lane for full-grid passes (x fastest, 64 voxels of a row per
// wavefront so every neighbour load of the wave is one coalesced 512 B row),
// grid-stride loops over device-resident lists for sparse passes (the list
// length never leaves the device), wavefront ballot + one atomic per wave for
// list appends, and an order-preserving int64 atomicMax for delta proposals.
#include "engine.h"

namespace ez {

__constant__ int c_dx[kNbr], c_dy[kNbr], c_dz[kNbr];
// The same stencil as compile-time tables, so unrolled loops get immediates
// (src/ExtreM_cuda.cu:155-162).
__device__ constexpr int kSX[kNbr] = {1, -1, 0, 0, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1};
__device__ constexpr int kSY[kNbr] = {0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1};
__device__ constexpr int kSZ[kNbr] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1};
__constant__ int c_cls_key[kNumCls];
__constant__ int c_lut_off[kNumCls];

void upload_constants(const int* dx, const int* dy, const int* dz,
                      const int* cls_key, const int* lut_off) {
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dx), dx, sizeof(int) * kNbr));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dy), dy, sizeof(int) * kNbr));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dz), dz, sizeof(int) * kNbr));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_cls_key), cls_key, sizeof(int) * kNumCls));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_lut_off), lut_off, sizeof(int) * kNumCls));
}

// ------------------------------------------------------------ helpers ----
__device__ __forceinline__ long long okey(double d) {
  long long b = __double_as_longlong(d);
  return b >= 0 ? b : (b ^ 0x7fffffffffffffffLL);
}
__device__ __forceinline__ double okey_inv(long long k) {
  return __longlong_as_double(k >= 0 ? k : (k ^ 0x7fffffffffffffffLL));
}

// (value, id) strict orders, isless / islarger of src/ExtreM.cpp:577-583.
__device__ __forceinline__ bool larger(double fu, int u, double fv, int v) {
  return fu > fv || (fu == fv && u > v);
}
__device__ __forceinline__ bool smaller(double fu, int u, double fv, int v) {
  return fu < fv || (fu == fv && u < v);
}

__device__ __forceinline__ void coords(const Geo& g, int v, int& x, int& y, int& z) {
  z = v / g.WH;
  int r = v - z * g.WH;
  y = r / g.W;
  x = r - y * g.W;
}

// Boundary class: per axis 0 interior, 1 at 0, 2 at max, 3 axis of length 1.
__device__ __forceinline__ int axis_state(int c, int n) {
  return n == 1 ? 3 : (c == 0 ? 1 : (c == n - 1 ? 2 : 0));
}
__device__ __forceinline__ int class_of(const Geo& g, int x, int y, int z) {
  if (g.planar) return kPlanarCls + axis_state(x, g.W) + 4 * axis_state(y, g.H);
  return axis_state(x, g.W) + 4 * axis_state(y, g.H) + 16 * axis_state(z, g.D);
}

// Appends v to list when pred; all lanes of the wave must call it.  A wave
// with no appending lane does nothing, so rare appends cost one ballot.
template <typename T>
__device__ __forceinline__ void wave_append(T* list, int* counter, bool pred, T v) {
  unsigned long long m = __ballot(pred);
  if (m == 0) return;
  int lane = __lane_id();
  int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

// Wavefront-collective append of cnt entries per lane to the list Q whose
// length is the LDS counter qn (one LDS atomic per wavefront); returns the
// lane's first index.  All lanes of the wave must call.
__device__ __forceinline__ int wave_reserve(int* qn, int cnt) {
  int pre = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(pre, o);
    if (__lane_id() >= o) pre += y;
  }
  const int total = __shfl(pre, 63);
  int base = 0;
  if (__lane_id() == 63 && total) base = atomicAdd(qn, total);
  return __shfl(base, 63) + pre - cnt;
}

// Blocks are dispatched round-robin over the 8 XCDs, each with its own L2.
// Returns a logical block index such that every XCD walks one contiguous
// range of tiles (so z +- 1 neighbour rows are in its own L2).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  int q = nb >> 3, r = nb & 7, x = b & 7, i = b >> 3;
  return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// Lower-star mask (direction space), compact upper config and the vertex type
// of classifyVertex_CUDA, src/ExtreM_cuda.cu:3796-3853.
__device__ __forceinline__ int classify_at(const Geo& g, const uint8_t* lut,
                                           const double* f, int v, int cls,
                                           unsigned* lower_out) {
  int key = c_cls_key[cls];
  double c = f[v];
  unsigned lm = 0, cfg = 0;
  int k = 0, lc = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (!((key >> d) & 1)) continue;
    int r = v + g.off[d];
    double fr = f[r];
    if (fr < c || (fr == c && r < v)) {
      lm |= 1u << d;
      lc++;
    } else {
      cfg |= 1u << k;
    }
    k++;
  }
  *lower_out = lm;
  // the GPU path's shortcuts (src/ExtreM_cuda.cu:3846-3853); the plane
  // adjacency keeps classifyVertex1's rule, whose link may be disconnected
  // (src/ExtreM.cpp:880-886: a vertex with two upper link components and no
  // lower one is a 2-saddle), so its table holds every configuration
  if (cls < kPlanarCls) {
    if (lc == k) return 4;   // no upper neighbour
    if (lc == 0) return 0;   // no lower neighbour
  }
  return lut[c_lut_off[cls] + cfg];
}

// Type of a known lower mask (LUT lookup only).
__device__ __forceinline__ int type_of_mask(const uint8_t* lut, int cls, unsigned lm) {
  const unsigned key = c_cls_key[cls];
  unsigned cfg = 0;
  int k = 0, lc = 0;
  if (key == (1u << kNbr) - 1) {  // interior: the upper config is the complement
    cfg = ~lm & key;
    k = kNbr;
    lc = __popc(lm & key);
  } else {
#pragma unroll
    for (int d = 0; d < kNbr; d++) {
      if (!((key >> d) & 1)) continue;
      if ((lm >> d) & 1) lc++;
      else cfg |= 1u << k;
      k++;
    }
  }
  if (cls < kPlanarCls) {  // as classify_at
    if (lc == k) return 4;
    if (lc == 0) return 0;
  }
  return lut[c_lut_off[cls] + cfg];
}

__device__ __forceinline__ unsigned lower_mask_at(const Geo& g, const double* f, int v, int cls) {
  int key = c_cls_key[cls];
  double c = f[v];
  unsigned lm = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (!((key >> d) & 1)) continue;
    int r = v + g.off[d];
    double fr = f[r];
    if (fr < c || (fr == c && r < v)) lm |= 1u << d;
  }
  return lm;
}

__device__ __forceinline__ bool is_saddle_type(int t) { return t >= 1 && t <= 3; }

// Whether v is listed by get_false_criticle_points (src/ExtreM.cpp:1324-1383).
__device__ __forceinline__ bool mismatch_at(const St& s, int v) {
  int x, y, z;
  coords(s.g, v, x, y, z);
  int cls = class_of(s.g, x, y, z);
  unsigned dl = lower_mask_at(s.g, s.decp, v, cls);
  unsigned ol = s.olow[v];
  if (dl == ol) return false;   // same star, hence same type
  int ot = s.otype[v];
  int dt = type_of_mask(s.lut, cls, dl);
  return dt != ot || is_saddle_type(ot);
}

// A list append staged in LDS: each push costs one LDS atomic; flush() moves
// the block's entries to the global list with one global atomic.  Entries
// beyond K go straight to the global list.  Every thread of the block must
// reach flush().
template <class T, int K>
struct StagedList {
  T* buf;
  int* n;
  int* base;
  T* glist;
  int* gcount;
  // Callable from divergent code: the active lanes of the wavefront share
  // one LDS atomic (and, once the block's buffer is full, one global atomic).
  __device__ __forceinline__ void push(T v) {
    unsigned long long m = __ballot(1);
    int lane = __lane_id();
    int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n, __popcll(m));
    base = __shfl(base, leader);
    int i = base + __popcll(m & ((1ull << lane) - 1ull));
    if (i < K) {
      buf[i] = v;
      return;
    }
    unsigned long long o = __ballot(1);  // the lanes past the buffer
    int ol = __ffsll((long long)o) - 1;
    int ob = 0;
    if (lane == ol) ob = atomicAdd(gcount, __popcll(o));
    ob = __shfl(ob, ol);
    glist[ob + __popcll(o & ((1ull << lane) - 1ull))] = v;
  }
  __device__ __forceinline__ void flush() {
    __syncthreads();
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nt = blockDim.x * blockDim.y;
    const int cnt = *n < K ? *n : K;
    if (tid == 0) *base = cnt ? atomicAdd(gcount, cnt) : 0;
    __syncthreads();
    for (int i = tid; i < cnt; i += nt) glist[*base + i] = buf[i];
  }
  // Moves the buffer to the global list and empties it when fewer than room
  // entries are free, so the pushes that follow stay in LDS.  Every thread of
  // the block calls it right after a barrier, with no push in flight (so all
  // read the same count).  room counts as at most half the buffer: a round
  // that could push more may spill its excess with a global atomic per
  // wavefront, but a small list is not moved out after every round.
  __device__ __forceinline__ void drain(int room) {
    const int cnt = *n < K ? *n : K;
    if (cnt == 0 || cnt + (room < K / 2 ? room : K / 2) <= K) return;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    const int nt = blockDim.x * blockDim.y;
    if (tid == 0) *base = atomicAdd(gcount, cnt);
    __syncthreads();
    for (int i = tid; i < cnt; i += nt) glist[*base + i] = buf[i];
    __syncthreads();
    if (tid == 0) *n = 0;
    __syncthreads();
  }
};
#define STAGED_LIST(name, T, K, glist, gcount)                         \
  __shared__ T name##_buf[K];                                          \
  __shared__ int name##_n, name##_base;                                \
  if (threadIdx.x == 0 && threadIdx.y == 0) name##_n = 0;              \
  __syncthreads();                                                     \
  StagedList<T, K> name{name##_buf, &name##_n, &name##_base, (glist), (gcount)}

// A counter staged in LDS (one global atomic per block at flush()).
struct StagedCount {
  int* n;
  int* gcount;
  // Adds one per calling lane; callable from divergent code (one LDS atomic
  // per wavefront).
  __device__ __forceinline__ void inc() {
    unsigned long long m = __ballot(1);
    if (__lane_id() == __ffsll((long long)m) - 1) atomicAdd(n, __popcll(m));
  }
  __device__ __forceinline__ void flush() {
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0 && *n) atomicAdd(gcount, *n);
  }
};
#define STAGED_COUNT(name, gcount)                                     \
  __shared__ int name##_n;                                             \
  if (threadIdx.x == 0 && threadIdx.y == 0) name##_n = 0;              \
  __syncthreads();                                                     \
  StagedCount name{&name##_n, (gcount)}

constexpr int kStage = 1024;
using Touched = StagedList<int, kStage>;

// swap + the caller's `d > oldValue` gate, sequentially a max (D5).  The
// first proposal above the sentinel lists the vertex as touched.
template <class TL>
__device__ __forceinline__ void propose(const St& s, int t, double d, TL& tl) {
  // fixed-step rule: every proposal is -xi/q; the float32 fork's fixes are
  // -bound / (1 << 2) (src/ExtreM_cuda_f.cu:2703, 2749, 2804)
  if (s.step_q || s.f32) d = s.step;
  long long k = okey(d);
  if (k <= s.sent) return;
  long long old = atomicMax(&s.delta[t], k);
  if (old == s.sent) tl.push(t);
}
// The star fixes of get_false_stars, (orig - xi) - decp of a vertex: with the
// float32 rule set in float32 (src/ExtreM_cuda_f.cu:2583-2600), with the
// fixed-step rule the step.
__device__ __forceinline__ double star_fix(const St& s, double o, double x) {
  if (s.step_q) return s.step;
  if (s.f32) return (double)(((float)o - (float)s.xi) - (float)x);
  return (o - s.xi) - x;
}
template <class TL>
__device__ __forceinline__ void propose_star(const St& s, int t, double d, TL& tl) {
  long long k = okey(d);
  if (k <= s.sent) return;
  long long old = atomicMax(&s.delta[t], k);
  if (old == s.sent) tl.push(t);
}
__device__ __forceinline__ double lower_mid(const St& s, int v) {
  return ((s.orig[v] - s.xi) + s.decp[v]) / 2.0 - s.decp[v];
}
__device__ __forceinline__ double raise_mid(const St& s, int v) {
  return ((s.orig[v] + s.xi) + s.decp[v]) / 2.0 - s.decp[v];
}

#define GRID_STRIDE(i, n) \
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

// GRID_STRIDE in block-uniform rounds for loops that append to staged lists:
// body(i) for this thread's item of a round (returning early is its
// `continue`), then every list is emptied when fewer entries are free than a
// round can push (per: the most one item pushes to a list), so a long input
// never appends with a global atomic per wavefront on one counter.  Every
// thread of the (1-D) block must call it.
template <class F, class... Ls>
__device__ __forceinline__ void block_rounds(long long n, int per, F&& body, Ls&... ls) {
  for (long long k0 = blockIdx.x * (long long)blockDim.x; k0 < n;
       k0 += (long long)gridDim.x * blockDim.x) {
    if (k0 + threadIdx.x < n) body(k0 + threadIdx.x);
    __syncthreads();
    (ls.drain(per * (int)blockDim.x), ...);
  }
}

// ---------------------------------------------------------- utilities ----
__global__ void k_fill_i64(long long* p, long long v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_i32(int* p, int v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void k_ctr_set(int* ctr, int idx, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr[idx] = v;
}
__global__ void k_ctr_set2(int* ctr, int a, int b, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ctr[a] = v;
    ctr[b] = v;
  }
}
// The counters into coherent host memory, then seq after them (system-scope
// release), on which the host waits.
__global__ void k_ctr_export(const int* ctr, int* host, int n, int seq) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) host[i] = ctr[i];
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(&host[n], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
__global__ void k_f32_to_f64(const float* in, double* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (double)in[i];
}
// Sets ctr[C_FAIL] when any value is not finite.
// Non-finite input fails the call; with the float32 rule set, so does a
// value float32 cannot hold exactly (the fork keeps its fields in float).
__global__ void k_check_finite(const double* a, long long n, int* ctr, int f32) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    if (!isfinite(a[i]) || (f32 && (double)(float)a[i] != a[i])) ctr[C_FAIL] = 1;
}

// ---------------------------------------------------- classification ----
// Row blocks: 64x4 voxels per 256-thread block (one 64-voxel row per
// wavefront), 1-D grid in XCD-aware order.
struct RowPos {
  int x, y, z;
  bool in;
};
__device__ __forceinline__ RowPos row_pos_at(const Geo& g, int L) {
  const int nbx = (g.W + 63) / 64, nby = (g.H + 3) / 4;
  RowPos p;
  p.x = (L % nbx) * 64 + threadIdx.x;
  p.y = ((L / nbx) % nby) * 4 + threadIdx.y;
  p.z = L / (nbx * nby);
  p.in = p.x < g.W && p.y < g.H;
  return p;
}
__device__ __forceinline__ RowPos row_pos(const Geo& g) {
  return row_pos_at(g, xcd_remap(blockIdx.x, gridDim.x));
}

// Full-grid classification of a field (classifyAllVertices,
// src/ExtreM.cpp:1088-1097).
__global__ void __launch_bounds__(256) k_classify(Geo g, const uint8_t* lut, const double* f,
                                                  uint8_t* type, uint16_t* lower) {
  RowPos p = row_pos(g);
  if (!p.in) return;
  int v = p.x + g.W * (p.y + g.H * p.z);
  unsigned lm;
  int t = classify_at(g, lut, f, v, class_of(g, p.x, p.y, p.z), &lm);
  type[v] = (uint8_t)t;
  lower[v] = (uint16_t)lm;
}

// Re-check after edits: every list entry (and, with nbrs, its 14 neighbours)
// is evaluated once per epoch; mismatches go to the frontier.  Statuses of
// vertices outside the edited neighbourhoods cannot change, so old frontier
// entries plus edited neighbourhoods give the full get_false_criticle_points
// result.  With clear_dirty the list is the dirty list (length ctr[lctr]): its
// flags are cleared and the other dirty length is set to zero for the next
// producer.
__global__ void __launch_bounds__(256) k_check_list(St s, const int* list, int lctr,
                                                    int with_nbrs, int* front, int fctr,
                                                    int epoch, int clear_dirty, int zero_idx) {
  STAGED_LIST(fl, int, kStage, front, &s.ctr[fctr]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (clear_dirty) s.ctr[lctr == C_DIRTY ? C_DIRTY2 : C_DIRTY] = 0;
    if (zero_idx >= 0) s.ctr[zero_idx] = 0;  // a length this launch does not read
  }
  // one lane per (entry, candidate): candidate 0 is the entry itself,
  // candidate 1 + d its neighbour in direction d
  const int per = with_nbrs ? kNbr + 1 : 1;
  const long long n = (long long)s.ctr[lctr] * per;
  auto one = [&](long long k) {
    const int i = (int)(k / per), j = (int)(k - (long long)i * per);
    int u = list[i];
    if (clear_dirty && j == 0) s.dflag[u] = 0;
    if (j > 0) {
      int x, y, z;
      coords(s.g, u, x, y, z);
      if (!((c_cls_key[class_of(s.g, x, y, z)] >> (j - 1)) & 1)) return;
      u += s.g.off[j - 1];
    }
    if (atomicExch(&s.stamp[u], epoch) == epoch) return;
    if (mismatch_at(s, u)) fl.push(u);
  };
  // block-uniform rounds of at most one push per thread, the buffer emptied
  // between them (a long list never appends with a global atomic per
  // wavefront on one counter)
  for (long long k0 = blockIdx.x * (long long)blockDim.x; k0 < n;
       k0 += (long long)gridDim.x * blockDim.x) {
    if (k0 + threadIdx.x < n) one(k0 + threadIdx.x);
    __syncthreads();
    fl.drain(256);
  }
  fl.flush();
}

// The two re-checks of a c_loops round in one launch: lanes [0, nf) take the
// last frontier's entries (nf passed by value, so its length can be cleared
// here: zero_idx), the rest one (dirty vertex, candidate) each as
// k_check_list with nbrs and clear_dirty.  One epoch stamp serves both, as
// for the two launches it replaces.
__global__ void __launch_bounds__(256) k_check_two(St s, const int* flist, int nf, int* front,
                                                   int fctr, int epoch, int zero_idx) {
  STAGED_LIST(fl, int, kStage, front, &s.ctr[fctr]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.ctr[s.dctr == C_DIRTY ? C_DIRTY2 : C_DIRTY] = 0;
    s.ctr[zero_idx] = 0;
  }
  const long long n = nf + (long long)s.ctr[s.dctr] * (kNbr + 1);
  auto one = [&](long long k) {
    int u;
    if (k < nf) {
      u = flist[k];
    } else {
      const long long q = k - nf;
      const int i = (int)(q / (kNbr + 1)), j = (int)(q - (long long)i * (kNbr + 1));
      u = s.dirty[i];
      if (j == 0) {
        s.dflag[u] = 0;
      } else {
        int x, y, z;
        coords(s.g, u, x, y, z);
        if (!((c_cls_key[class_of(s.g, x, y, z)] >> (j - 1)) & 1)) return;
        u += s.g.off[j - 1];
      }
    }
    if (atomicExch(&s.stamp[u], epoch) == epoch) return;
    if (mismatch_at(s, u)) fl.push(u);
  };
  // block-uniform rounds, the buffer emptied between them (as k_check_list)
  for (long long k0 = blockIdx.x * (long long)blockDim.x; k0 < n;
       k0 += (long long)gridDim.x * blockDim.x) {
    if (k0 + threadIdx.x < n) one(k0 + threadIdx.x);
    __syncthreads();
    fl.drain(256);
  }
  fl.flush();
}

// Clears the dirty flags of the listed vertices; the other dirty length (and
// zero_idx, when given) is set to zero for the next producer.
__global__ void k_clear_dirty(St s, int zero_idx) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.ctr[s.dctr == C_DIRTY ? C_DIRTY2 : C_DIRTY] = 0;
    if (zero_idx >= 0) s.ctr[zero_idx] = 0;
  }
  int n = s.ctr[s.dctr];
  GRID_STRIDE(i, n) s.dflag[s.dirty[i]] = 0;
}

// ------------------------------------------------- c_loops proposals ----
// c_loop (src/ExtreM.cpp:1152-1236) for false maxima / minima and
// get_false_stars (src/ExtreM.cpp:1412-1445) for saddle-list entries.
// A frontier entry makes at most 15 proposals (itself or its 14 neighbours):
// block-uniform rounds with the touched buffer emptied whenever a round
// could overfill it, so a dense frontier never appends with a global atomic
// per wavefront on one counter.
constexpr int kProposeStage = 4096;
__global__ void __launch_bounds__(256) k_propose_c(St s, const int* front, int fctr) {
  STAGED_LIST(tl, int, kProposeStage, s.touched, &s.ctr[s.tctr]);
  const int n = s.ctr[fctr];
  for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    const int i = i0 + threadIdx.x;
    if (i < n) {
      int v = front[i];
      int x, y, z;
      coords(s.g, v, x, y, z);
      int cls = class_of(s.g, x, y, z);
      int key = c_cls_key[cls];
      unsigned dl = lower_mask_at(s.g, s.decp, v, cls);
      unsigned ol = s.olow[v];
      int dt = type_of_mask(s.lut, cls, dl);
      int ot = s.otype[v];
      if ((dt == 4) != (ot == 4)) {  // all_max, direction 0
        if (ot != 4) {
          propose(s, v, lower_mid(s, v), tl);
        } else {
          int largest = v;
          double fl = s.decp[v];
          for (int d = 0; d < kNbr; d++) {
            if (!((key >> d) & 1)) continue;
            int r = v + s.g.off[d];
            double fr = s.decp[r];
            if (larger(fr, r, fl, largest)) {
              largest = r;
              fl = fr;
            }
          }
          double fv = s.decp[v];
          if (!(fv > fl || (fv == fl && v > largest))) propose(s, largest, lower_mid(s, largest), tl);
        }
      }
      if ((dt == 0) != (ot == 0)) {  // all_min, direction 1
        if (ot != 0) {
          int smallest = v;
          double fo = s.orig[v];
          for (int d = 0; d < kNbr; d++) {
            if (!((key >> d) & 1)) continue;
            int r = v + s.g.off[d];
            double fr = s.orig[r];
            if (smaller(fr, r, fo, smallest)) {
              smallest = r;
              fo = fr;
            }
          }
          double dv = s.decp[v], ds = s.decp[smallest];
          if (!(dv > ds || (dv == ds && v > smallest)))
            propose(s, smallest, lower_mid(s, smallest), tl);
        } else {
          propose(s, v, lower_mid(s, v), tl);
        }
      }
      bool sad = (dt != ot) ? (is_saddle_type(dt) || is_saddle_type(ot))
                            : (is_saddle_type(ot) && dl != ol);
      if (sad && dl != ol) {  // get_false_stars
        // propose() for every star neighbour, with all loads and atomics of the
        // star issued before any result is used
        const unsigned a = ol & ~dl;
        double po[kNbr], pd[kNbr];
  #pragma unroll
        for (int d = 0; d < kNbr; d++) {
          const bool on = (a >> d) & 1;
          po[d] = on ? s.orig[v + s.g.off[d]] : 0.0;
          pd[d] = on ? s.decp[v + s.g.off[d]] : 0.0;
        }
        long long old[kNbr];
        unsigned sent = 0;  // atomics issued
  #pragma unroll
        for (int d = 0; d < kNbr; d++) {
          old[d] = 0;
          if (!((a >> d) & 1)) continue;
          const long long k = okey(star_fix(s, po[d], pd[d]));
          if (k <= s.sent) continue;
          old[d] = atomicMax(&s.delta[v + s.g.off[d]], k);
          sent |= 1u << d;
        }
  #pragma unroll
        for (int d = 0; d < kNbr; d++)
          if (((sent >> d) & 1) && old[d] == s.sent) tl.push(v + s.g.off[d]);
        if (dl & ~ol) propose_star(s, v, star_fix(s, s.orig[v], s.decp[v]), tl);
      }
    }
    __syncthreads();
    tl.drain(256 * (kNbr + 1));
  }
  tl.flush();
}

// apply_delta, split (src/ExtreM.cpp:1486-1503), or with the fixed-step rule
// the edit_type 0 branch of applyDeltaBuffer, over the touched list only;
// resets the delta slot and records changed vertices as dirty (for the next
// false-critical-point check) and as changed since the last CP pass of decp.
// The other touched length (and zero_idx, when given) is set to zero for the
// next producer.
__global__ void __launch_bounds__(256) k_apply(St s, int zero_idx) {
  __shared__ int changed;  // one store of C_ANY_CHANGE per block, not one per lane
  if (threadIdx.x == 0) changed = 0;
  STAGED_LIST(dl, int, kStage, s.dirty, &s.ctr[s.dctr]);
  STAGED_LIST(cl, int, kStage, s.lchg, &s.ctr[C_LCHG]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.ctr[s.tctr == C_TOUCHED ? C_TOUCHED2 : C_TOUCHED] = 0;
    if (zero_idx >= 0) s.ctr[zero_idx] = 0;
  }
  const int n = s.ctr[s.tctr];
  // block-uniform rounds: a round pushes at most one entry per thread to
  // each list, and the buffers are emptied between rounds, so a long touched
  // list never appends with a global atomic per wavefront on one counter
  for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
    const int i = i0 + threadIdx.x;
    if (i < n) {
      int t = s.touched[i];
      double dd = okey_inv(s.delta[t]);
      s.delta[t] = s.sent;
      double o = s.orig[t], x = s.decp[t];
      int c = s.cnt[t];
      double nx;
      if (s.f32) {  // applyDeltaBuffer of the float32 fork (src/ExtreM_cuda_f.cu:2854-2866)
        const float df = (float)dd, xf = (float)x, of = (float)o, bf = (float)s.xi;
        if (fabsf(df) > 1e-6 && c < 6 && fabsf(of - (xf + df)) < bf) {
          nx = (double)(xf + df);
          c += 1;
        } else {
          nx = (double)(of - bf);
          c = 7;
        }
      } else if (s.step_q) {  // applyDeltaBuffer, edit_type 0 (src/ExtreM_cuda.cu:3065-3076)
        if (c < s.step_q - 1 && fabs(o - (x + dd)) <= s.xi) {
          nx = x + dd;
          c += 1;
        } else {
          nx = o - s.xi;
          c = s.step_q;
        }
      } else if (fabs(dd) > 1e-15 && c < 5 && fabs(o - x - dd) <= s.xi) {
        nx = x + dd;
        c += 1;
      } else {
        nx = o - s.xi;
        c = 6;
      }
      s.decp[t] = nx;
      s.cnt[t] = (uint8_t)c;
      if (nx != x) {
        changed = 1;
        if (!s.dflag[t]) {
          s.dflag[t] = 1;
          dl.push(t);
        }
        if (!s.lflag[t]) {  // one touched-list entry per vertex: no other lane sets it
          s.lflag[t] = 1;
          cl.push(t);
        }
      }
    }
    __syncthreads();
    dl.drain(256);
    cl.drain(256);
  }
  dl.flush();
  cl.flush();
  if (threadIdx.x == 0 && changed) s.ctr[C_ANY_CHANGE] = 1;  // (after the flushes' barriers)
}

// ----------------------------------------------------- manifold labels ----
// Tiles of 64x4x8 voxels (kLabelTileZ planes), marched along z through a ring
// of three LDS planes (66x6 values each, halo included; the next plane is
// prefetched into registers while the current one is processed).  Steepest
// ascent / descent pointers (first loops of src/ExtreM.cpp:146-157 and
// 504-515) are computed from LDS and resolved inside the tile by walking them
// in LDS; a pointer then names the root or the first vertex outside the tile
// on the steepest path.  With lower != nullptr the lower-star mask of every
// voxel is written too (with TYPE also its vertex type from the LUT, the
// classification of src/ExtreM_cuda.cu:3796-3853), with pdir != nullptr its
// steepest directions ((up + 1) | (down + 1) << 4, 0 for the voxel itself),
// the state the incremental CP pass starts from.
//
// FAST (fields with W >= 3 on the stencil): the 15 candidate ids (self + 14
// neighbours) have the fixed order of their offsets
//   d9 < d13 < d5 < d11 < d3 < d7 < d1 < self < d0 < d6 < d2 < d10 < d4 < d12 < d8
// for every voxel whose neighbours exist, so each (value, id) comparison of
// src/ExtreM.cpp:153 / 511 is one compare.  A neighbour outside the field
// reads NaN in the ring, for which every comparison is false: it is neither
// in the lower star nor ever a steepest choice, so boundary voxels take the
// same straight-line code.  The scans start from the voxel itself, visit the
// smaller ids in descending order (a tie keeps the larger id already held on
// the way up, takes the smaller new id on the way down), then the larger ids
// in ascending order (a tie takes the larger new id up, keeps the smaller
// one down).  Otherwise (W < 3, or the plane adjacency) each voxel loops over
// the directions its class key holds, with explicit id tie-breaks.
constexpr int kLTX = 64, kLTY = 4, kLTZ = kLabelTileZ, kLTN = kLTX * kLTY * kLTZ;
constexpr int kExit = (int)0x80000000u;  // label bit: a pointer out of the tile, not a root
constexpr int kPX = kLTX + 2, kPY = kLTY + 2, kPN = kPX * kPY;
constexpr int kPPer = (kPN + 255) / 256;
// stencil directions leaving a tile through its -x / +x / -y / +y / -z / +z face
constexpr unsigned kFaceXN = 0x1442u, kFaceXP = 0x2881u, kFaceYN = 0x2288u, kFaceYP = 0x1144u,
                   kFaceZN = 0x2A20u, kFaceZP = 0x1510u;

template <bool FAST, bool TYPE>
__global__ void __launch_bounds__(256) k_labels_tile(Geo g, const double* f, int* ds, int* as,
                                                     uint16_t* lower, uint8_t* pdir,
                                                     const uint8_t* lut, uint8_t* type,
                                                     Presnap ps) {
  __shared__ double ring[3][kPN];
  __shared__ int sup[kLTN];
  __shared__ int sdn[kLTN];
  __shared__ int s_goff[kNbr], s_loff[kNbr];  // grid and in-tile offset of each direction
  const int ntx = (g.W + kLTX - 1) / kLTX, ntz = (g.D + kLTZ - 1) / kLTZ;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  // z fastest: the consecutive tiles of an XCD stack along z and share their
  // halo planes in that XCD's L2
  const int col = L / ntz;
  const int x0 = (col % ntx) * kLTX, y0 = (col / ntx) * kLTY, z0 = (L % ntz) * kLTZ;
  const int lx = threadIdx.x, ly = threadIdx.y, tid = lx + 64 * ly;
  if (tid == 0) {
#pragma unroll
    for (int d = 0; d < kNbr; d++) {
      s_goff[d] = g.off[d];
      s_loff[d] = kSX[d] + kLTX * (kSY[d] + kLTY * kSZ[d]);
    }
  }
  const double miss = __longlong_as_double(0x7ff8000000000000LL);  // quiet NaN
  const int lut0 = TYPE ? c_lut_off[0] : 0;  // the interior class's LUT segment
  // this thread's halo slots of a plane: fixed offsets within every plane
  int foff[kPPer];
  bool fin[kPPer];
#pragma unroll
  for (int j = 0; j < kPPer; j++) {
    const int i = tid + 256 * j;
    const int px = i % kPX, py = i / kPX, gx = x0 + px - 1, gy = y0 + py - 1;
    fin[j] = i < kPN && gx >= 0 && gx < g.W && gy >= 0 && gy < g.H;
    foff[j] = gx + g.W * gy;
  }
  double pre[kPPer];
  auto fetch = [&](int z) {
    const bool zin = z >= 0 && z < g.D;
    const long long zo = (long long)z * g.WH;
#pragma unroll
    for (int j = 0; j < kPPer; j++)
      pre[j] = (fin[j] && zin) ? f[foff[j] + zo] : (FAST ? miss : 0.0);
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int j = 0; j < kPPer; j++) {
      const int i = tid + 256 * j;
      if (i < kPN) ring[slot][i] = pre[j];
    }
  };
  // directions that stay inside the tile in x and y (z: per plane below)
  unsigned in_xy = (1u << kNbr) - 1;
  if (lx == 0) in_xy &= ~kFaceXN;
  if (lx == kLTX - 1) in_xy &= ~kFaceXP;
  if (ly == 0) in_xy &= ~kFaceYN;
  if (ly == kLTY - 1) in_xy &= ~kFaceYP;
  const int x = x0 + lx, y = y0 + ly;
  const bool xy_in = x < g.W && y < g.H;
  fetch(z0 - 1);
  store(0);
  fetch(z0);
  store(1);
  fetch(z0 + 1);
#pragma unroll
  for (int lz = 0; lz < kLTZ; lz++) {
    store((lz + 2) % 3);
    __syncthreads();
    if (lz + 1 < kLTZ) fetch(z0 + lz + 2);
    const double* P0 = ring[lz % 3];        // z - 1
    const double* P1 = ring[(lz + 1) % 3];  // z
    const double* P2 = ring[(lz + 2) % 3];  // z + 1
    const int z = z0 + lz;
    const int li = lx + kLTX * (ly + kLTY * lz);
    int pu = li, pd = li;
    if (xy_in && z < g.D) {
      const int v = x + g.W * (y + g.H * z);
      const int c0 = (lx + 1) + kPX * (ly + 1);
      const double c = P1[c0];
      int du = -1, dd = -1, cls = 0;
      unsigned lm = 0;
      if (FAST) {
        const double n0 = P1[c0 + 1], n1 = P1[c0 - 1], n2 = P1[c0 + kPX], n3 = P1[c0 - kPX];
        const double n4 = P2[c0], n5 = P0[c0], n6 = P1[c0 - 1 + kPX], n7 = P1[c0 + 1 - kPX];
        const double n8 = P2[c0 + kPX], n9 = P0[c0 - kPX], n10 = P2[c0 - 1], n11 = P0[c0 + 1];
        const double n12 = P2[c0 - 1 + kPX], n13 = P0[c0 + 1 - kPX];
        // lower: neighbours with a smaller id (odd directions) count on ties
        lm = (unsigned)(n0 < c) | (unsigned)(n1 <= c) << 1 | (unsigned)(n2 < c) << 2 |
             (unsigned)(n3 <= c) << 3 | (unsigned)(n4 < c) << 4 | (unsigned)(n5 <= c) << 5 |
             (unsigned)(n6 < c) << 6 | (unsigned)(n7 <= c) << 7 | (unsigned)(n8 < c) << 8 |
             (unsigned)(n9 <= c) << 9 | (unsigned)(n10 < c) << 10 | (unsigned)(n11 <= c) << 11 |
             (unsigned)(n12 < c) << 12 | (unsigned)(n13 <= c) << 13;
        double bu = c, bd = c;
#define EZ_LO(val, dir)   \
  if ((val) > bu) {       \
    bu = (val);           \
    du = (dir);           \
  }                       \
  if ((val) <= bd) {      \
    bd = (val);           \
    dd = (dir);           \
  }
#define EZ_HI(val, dir)   \
  if ((val) >= bu) {      \
    bu = (val);           \
    du = (dir);           \
  }                       \
  if ((val) < bd) {       \
    bd = (val);           \
    dd = (dir);           \
  }
        EZ_LO(n1, 1) EZ_LO(n7, 7) EZ_LO(n3, 3) EZ_LO(n11, 11) EZ_LO(n5, 5) EZ_LO(n13, 13)
        EZ_LO(n9, 9)
        EZ_HI(n0, 0) EZ_HI(n6, 6) EZ_HI(n2, 2) EZ_HI(n10, 10) EZ_HI(n4, 4) EZ_HI(n12, 12)
        EZ_HI(n8, 8)
#undef EZ_LO
#undef EZ_HI
        if (TYPE) cls = class_of(g, x, y, z);
      } else {
        cls = class_of(g, x, y, z);
        const int key = c_cls_key[cls];
        int up = v, dn = v;
        double fu = c, fd = c;
#pragma unroll
        for (int d = 0; d < kNbr; d++) {
          if (!((key >> d) & 1)) continue;
          const double* P = kSZ[d] < 0 ? P0 : (kSZ[d] > 0 ? P2 : P1);
          double fr = P[c0 + kSX[d] + kPX * kSY[d]];
          int r = v + g.off[d];
          if (fr < c || (fr == c && r < v)) lm |= 1u << d;
          if (fu < fr || (fu == fr && up < r)) { up = r; fu = fr; du = d; }
          if (fd > fr || (fd == fr && dn > r)) { dn = r; fd = fr; dd = d; }
        }
      }
      if (lower) lower[v] = (uint16_t)lm;
      if (TYPE) {  // classifyVertex
        int t;
        if (cls == 0) {  // interior: the upper config is the complement of the lower mask
          const unsigned cfg = ~lm & ((1u << kNbr) - 1);
          t = cfg == 0 ? 4 : (cfg == (1u << kNbr) - 1 ? 0 : lut[lut0 + cfg]);
        } else {
          t = type_of_mask(lut, cls, lm);
        }
        type[v] = (uint8_t)t;
        // the pre-snap (main.cpp:74-75, 122-131): decp := orig - xi at every
        // saddle (types 1-3) and at the global minimum and maximum
        if (ps.decp && ((t >= 1 && t <= 3) || v == ps.ctr[C_GMIN] || v == ps.ctr[C_GMAX]))
          ps.decp[v] = ps.f32 ? (double)((float)c - (float)ps.xi) : c - ps.xi;
      }
      if (pdir) pdir[v] = (uint8_t)((du + 1) | ((dd + 1) << 4));
      unsigned inside = in_xy;
      if (lz == 0) inside &= ~kFaceZN;
      if (lz == kLTZ - 1) inside &= ~kFaceZP;
      if (du >= 0) pu = ((inside >> du) & 1) ? li + s_loff[du] : ~(v + s_goff[du]);
      if (dd >= 0) pd = ((inside >> dd) & 1) ? li + s_loff[dd] : ~(v + s_goff[dd]);
    }
    sup[li] = pu;
    sdn[li] = pd;
    __syncthreads();
  }
  // each voxel walks its in-tile path to a root or an exit (paths are short
  // on rough fields; a value read on the way is the next step, or, once its
  // own lane has finished, that path's end: either is an ancestor)
#pragma unroll
  for (int k = 0; k < kLTN / 256; k++) {
    const int li = tid + 256 * k;
    int c = li, p = sup[li];
    for (int step = 0; step < kLTN && p >= 0 && p != c; step++) {
      c = p;
      p = sup[c];
    }
    sup[li] = p >= 0 ? c : p;
    c = li;
    p = sdn[li];
    for (int step = 0; step < kLTN && p >= 0 && p != c; step++) {
      c = p;
      p = sdn[c];
    }
    sdn[li] = p >= 0 ? c : p;
  }
  __syncthreads();
  // a root inside the tile is final as written; a pointer out of the tile
  // carries the top bit, and only those are left to k_chase
  const int vb = x0 + g.W * (y0 + g.H * z0);
#pragma unroll
  for (int k = 0; k < kLTN / 256; k++) {
    const int li = tid + 256 * k;
    const int xx = li % kLTX, yy = (li / kLTX) % kLTY, zz = li / (kLTX * kLTY);
    if (x0 + xx >= g.W || y0 + yy >= g.H || z0 + zz >= g.D) continue;
    const int v = vb + xx + g.W * (yy + g.H * zz);
    int p = sup[li];
    ds[v] = p >= 0 ? vb + (p % kLTX) + g.W * ((p / kLTX) % kLTY + g.H * (p / (kLTX * kLTY)))
                   : kExit | ~p;
    p = sdn[li];
    as[v] = p >= 0 ? vb + (p % kLTX) + g.W * ((p / kLTX) % kLTY + g.H * (p / (kLTX * kLTY)))
                   : kExit | ~p;
  }
}
template __global__ void k_labels_tile<true, true>(Geo, const double*, int*, int*, uint16_t*,
                                                   uint8_t*, const uint8_t*, uint8_t*, Presnap);
template __global__ void k_labels_tile<true, false>(Geo, const double*, int*, int*, uint16_t*,
                                                    uint8_t*, const uint8_t*, uint8_t*, Presnap);
template __global__ void k_labels_tile<false, true>(Geo, const double*, int*, int*, uint16_t*,
                                                    uint8_t*, const uint8_t*, uint8_t*, Presnap);
template __global__ void k_labels_tile<false, false>(Geo, const double*, int*, int*, uint16_t*,
                                                     uint8_t*, const uint8_t*, uint8_t*, Presnap);

// Full-grid lower masks through the same LDS ring as k_labels_tile: a block
// marches one 64x4 column through zlen planes (z chunk zc of the column), so
// each decp value is read once per plane and the halo planes are shared.  A
// present direction d names a neighbour with a larger id for even d and a
// smaller one for odd d, so a tie counts as lower exactly for odd d
// (src/ExtreM.cpp:577-583).  MODE 1 (the one instantiated): the full
// false-critical-point scan of get_false_criticle_points (src/ExtreM.cpp:
// 1513-1519), appending to the frontier, one append per block; the
// original field's classification comes out of its label pass.
template <int MODE, bool FAST, int TY>
__global__ void __launch_bounds__(64 * TY) k_mask_cols(St s, const double* f, uint8_t* type,
                                                       uint16_t* lower, int* front, int fctr,
                                                       int nzc, int zlen) {
  constexpr int NT = 64 * TY, PN = kPX * (TY + 2), PPer = (PN + NT - 1) / NT;
  __shared__ double ring[3][PN];
  STAGED_LIST(fl, int, 2048, front, &s.ctr[fctr]);
  const Geo& g = s.g;
  const int ntx = (g.W + 63) / 64;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int col = L / nzc;
  const int x0 = (col % ntx) * 64, y0 = (col / ntx) * TY;
  const int z0 = (L % nzc) * zlen, z1 = min(g.D, z0 + zlen);
  const int lx = threadIdx.x, ly = threadIdx.y, tid = lx + 64 * ly;
  double pre[PPer];
  auto fetch = [&](int z) {
#pragma unroll
    for (int j = 0; j < PPer; j++) {
      int i = tid + NT * j;
      // FAST: a neighbour outside the field reads NaN, so every compare with it
      // is false and no class key is needed to leave it out
      double val = FAST ? __longlong_as_double(0x7ff8000000000000LL) : 0.0;
      if (i < PN) {
        int px = i % kPX, py = i / kPX, gx = x0 + px - 1, gy = y0 + py - 1;
        if (z >= 0 && z < g.D && gx >= 0 && gx < g.W && gy >= 0 && gy < g.H)
          val = f[gx + g.W * (gy + g.H * z)];
      }
      pre[j] = val;
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int j = 0; j < PPer; j++) {
      int i = tid + NT * j;
      if (i < PN) ring[slot][i] = pre[j];
    }
  };
  if (z0 < z1) {
    fetch(z0 - 1);
    store(0);
    fetch(z0);
    store(1);
    fetch(z0 + 1);
  }
  const int x = x0 + lx, y = y0 + ly;
  const int c0 = (lx + 1) + kPX * (ly + 1);
  for (int z = z0; z < z1; z++) {
    const int lz = z - z0;
    store((lz + 2) % 3);
    __syncthreads();
    // a plane pushes at most one entry per owned voxel: the buffer is emptied
    // before it could fill (a dense frontier, up to most of the grid, would
    // otherwise append with a global atomic per wavefront on one counter)
    fl.drain(NT);
    if (z + 2 < z1 + 1) fetch(z + 2);
    const double* P0 = ring[lz % 3];        // z - 1
    const double* P1 = ring[(lz + 1) % 3];  // z
    const double* P2 = ring[(lz + 2) % 3];  // z + 1
    if (x < g.W && y < g.H) {
      const int v = x + g.W * (y + g.H * z);
      const unsigned key = FAST ? (1u << kNbr) - 1 : c_cls_key[class_of(g, x, y, z)];
      const double c = P1[c0];
      unsigned lm = 0;
#pragma unroll
      for (int d = 0; d < kNbr; d++) {
        const double* P = kSZ[d] < 0 ? P0 : (kSZ[d] > 0 ? P2 : P1);
        const double fr = P[c0 + kSX[d] + kPX * kSY[d]];
        const bool lo = fr < c || ((d & 1) && fr == c);
        lm |= (unsigned)(((key >> d) & 1) && lo) << d;
      }
      if (lm != s.olow[v]) {
        const int ot = s.otype[v];
        if (type_of_mask(s.lut, class_of(g, x, y, z), lm) != ot || is_saddle_type(ot))
          fl.push(v);
      }
    }
    __syncthreads();
  }
  fl.flush();
}
template __global__ void k_mask_cols<1, true, 4>(St, const double*, uint8_t*, uint16_t*, int*,
                                                 int, int, int);
template __global__ void k_mask_cols<1, false, 4>(St, const double*, uint8_t*, uint16_t*, int*,
                                                  int, int, int);

// The full false-critical-point scan of decp (k_mask_cols<1>) fused with the
// c_loops proposals of the frontier it finds (k_propose_c over that list).
// Every proposal's value is a function of its target alone: (orig - xi) -
// decp for a star neighbour or the vertex itself (get_false_stars,
// src/ExtreM.cpp:1412-1445), the lower midpoint for the c_loop cases
// (src/ExtreM.cpp:1152-1236).  So each voxel gathers what its neighbours
// send to it, and no atomics are needed.
// A block owns a 64x4 column over zlen planes. Each plane is processed as follows:
//   - decp, with a halo of two, is held in an LDS ring.
//   - The outgoing proposals of every voxel of the tile and its halo of one
//     are stored in a second ring (the word below).
//   - Each owned voxel ORs what its 14 neighbours and itself send to it and
//     writes its delta key. A plain store suffices: the voxel is the only
//     writer, and delta holds the empty key between edit rounds.
//   - The voxel then joins the touched list, and frontier entries are pushed
//     as k_mask_cols<1> pushes them.
// The outgoing word:
//   bits 0-13  (orig - xi) - decp to neighbour d
//   bit 14     the same to the voxel itself
//   bits 15-28 the lower midpoint to neighbour d (the largest decp or the
//              smallest original neighbour)
//   bit 29     the lower midpoint to the voxel itself
//   bit 30     the voxel is a frontier entry
constexpr int kQX = kLTX + 4;  // decp plane width of k_scan_propose (halo 2)

// Outgoing proposals of one voxel from LDS planes of decp around it (Q0..Q2,
// centre qc) and its original lower mask / type.
// With FAST (the stencil adjacency) the ring holds NaN outside the field, so
// the lower mask needs no class key; the key is looked up only for the few
// voxels whose mask differs from the original one.
template <bool FAST>
__device__ __forceinline__ unsigned out_proposals(const St& s, const double* Q0, const double* Q1,
                                                  const double* Q2, int qc, int x, int y, int z,
                                                  unsigned ol, int ot) {
  const Geo& g = s.g;
  const int v = x + g.W * (y + g.H * z);
  const int cls0 = FAST ? 0 : class_of(g, x, y, z);
  const unsigned key0 = FAST ? (1u << kNbr) - 1 : c_cls_key[cls0];
  const double c = Q1[qc];
  double fr[kNbr];
  unsigned dl = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    const double* Q = kSZ[d] < 0 ? Q0 : (kSZ[d] > 0 ? Q2 : Q1);
    fr[d] = Q[qc + kSX[d] + kQX * kSY[d]];
    const bool lo = fr[d] < c || ((d & 1) && fr[d] == c);
    dl |= (unsigned)(((key0 >> d) & 1) && lo) << d;
  }
  if (dl == ol) return 0;
  const int cls = FAST ? class_of(g, x, y, z) : cls0;
  const unsigned key = FAST ? c_cls_key[cls] : key0;
  const int dt = type_of_mask(s.lut, cls, dl);
  if (!(dt != ot || is_saddle_type(ot))) return 0;
  unsigned w = 1u << 30;
  if ((dt == 4) != (ot == 4)) {  // all_max
    if (ot != 4) {
      w |= 1u << 29;
    } else {
      int largest = v, ld = -1;
      double fl = c;
#pragma unroll
      for (int d = 0; d < kNbr; d++) {
        if (!((key >> d) & 1)) continue;
        const int r = v + g.off[d];
        if (larger(fr[d], r, fl, largest)) {
          largest = r;
          fl = fr[d];
          ld = d;
        }
      }
      if (!(c > fl || (c == fl && v > largest))) w |= ld < 0 ? 1u << 29 : 1u << (15 + ld);
    }
  }
  if ((dt == 0) != (ot == 0)) {  // all_min
    if (ot != 0) {
      // the smallest original neighbour: every load issued before the scan
      double on[kNbr];
#pragma unroll
      for (int d = 0; d < kNbr; d++) on[d] = ((key >> d) & 1) ? s.orig[v + g.off[d]] : 0.0;
      int smallest = v;
      double fo = s.orig[v], ds = c;
      unsigned sd = 0;  // 1 + direction of the smallest, 0 for the voxel itself
#pragma unroll
      for (int d = 0; d < kNbr; d++) {
        if (!((key >> d) & 1)) continue;
        const int r = v + g.off[d];
        if (smaller(on[d], r, fo, smallest)) {
          smallest = r;
          fo = on[d];
          ds = fr[d];
          sd = d + 1;
        }
      }
      if (!(c > ds || (c == ds && v > smallest))) w |= sd ? 1u << (14 + sd) : 1u << 29;
    } else {
      w |= 1u << 29;
    }
  }
  const bool sad = (dt != ot) ? (is_saddle_type(dt) || is_saddle_type(ot))
                              : (is_saddle_type(ot) && dl != ol);
  if (sad) {  // get_false_stars (dl != ol here)
    w |= ol & ~dl;
    if (dl & ~ol) w |= 1u << 14;
  }
  return w;
}

// Pipeline (one barrier per step): step k stores decp plane z0 + k, computes
// the words of plane p = z0 - 1 + k and gathers plane q = z0 - 3 + k, whose
// three word planes were finished in earlier steps.  decp plane j sits in
// ring slot (j - z0 + 3) % 5, the words of plane j in sring slot
// (j - z0 + 1) % 4; each slot is rewritten only a barrier after its last
// reader.
// TY rows of 64 columns per block (64 * TY threads); the word planes have a
// halo of 1, the decp planes a halo of 2.
template <bool FAST, int TY>
__global__ void __launch_bounds__(64 * TY) k_scan_propose(St s, int* front, int fctr, int nzc,
                                                          int zlen) {
  constexpr int NT = 64 * TY;
  constexpr int PN = kPX * (TY + 2), QN = kQX * (TY + 4);
  constexpr int PPer = (PN + NT - 1) / NT, QPer = (QN + NT - 1) / NT;
  static_assert(PPer == 2, "two word-plane entries per thread");
  __shared__ double ring[5][QN];
  __shared__ unsigned sring[4][PN];
  STAGED_LIST(fl, int, 1024, front, &s.ctr[fctr]);
  STAGED_LIST(tl, int, kStage, s.touched, &s.ctr[s.tctr]);
  const Geo& g = s.g;
  const int ntx = (g.W + 63) / 64;
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int col = L / nzc;
  const int x0 = (col % ntx) * 64, y0 = (col / ntx) * TY;
  const int z0 = (L % nzc) * zlen, z1 = min(g.D, z0 + zlen);
  const int lx = threadIdx.x, ly = threadIdx.y, tid = lx + 64 * ly;
  const int x = x0 + lx, y = y0 + ly;
  const bool own_in = x < g.W && y < g.H;
  // prefetched a step ahead: a decp plane (halo 2), the original lower masks
  // and types of a word plane (halo 1), orig of the owned voxel of a gather plane
  double pd[QPer];
  unsigned pol[PPer];
  int pot[PPer];
  double pog = 0.0;
  auto fetch = [&](int z) {
#pragma unroll
    for (int j = 0; j < QPer; j++) {
      const int i = tid + NT * j;
      double a = FAST ? __longlong_as_double(0x7ff8000000000000LL) : 0.0;  // outside: NaN
      if (i < QN) {
        const int gx = x0 + i % kQX - 2, gy = y0 + i / kQX - 2;
        if (z >= 0 && z < g.D && gx >= 0 && gx < g.W && gy >= 0 && gy < g.H)
          a = s.decp[gx + g.W * (gy + g.H * z)];
      }
      pd[j] = a;
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int j = 0; j < QPer; j++) {
      const int i = tid + NT * j;
      if (i < QN) ring[slot][i] = pd[j];
    }
  };
  auto fetch_words = [&](int z) {
#pragma unroll
    for (int j = 0; j < PPer; j++) {
      const int i = tid + NT * j;
      unsigned a = 0;
      int b = 0;
      if (i < PN) {
        const int gx = x0 + i % kPX - 1, gy = y0 + i / kPX - 1;
        if (z >= 0 && z < g.D && gx >= 0 && gx < g.W && gy >= 0 && gy < g.H) {
          const int v = gx + g.W * (gy + g.H * z);
          a = s.olow[v];
          b = s.otype[v];
        }
      }
      pol[j] = a;
      pot[j] = b;
    }
  };
  auto fetch_orig = [&](int z) {
    pog = (own_in && z >= z0 && z < z1) ? s.orig[x + g.W * (y + g.H * z)] : 0.0;
  };
  const int nk = z0 < z1 ? z1 - z0 + 3 : 0;
  if (nk) {
    fetch(z0 - 2);
    store(1);
    fetch(z0 - 1);
    store(2);
    fetch(z0);
    fetch_words(z0 - 1);
  }
  for (int k = 0; k < nk; k++) {
    const int p = z0 - 1 + k, q = z0 - 3 + k;
    const bool do_a = p <= z1, do_b = k >= 3;
    if (do_a) store((k + 3) % 5);
    const unsigned ol0 = pol[0], ol1 = pol[1];
    const int ot0 = pot[0], ot1 = pot[1];
    const double og = pog;
    __syncthreads();
    // a step pushes at most one entry per owned voxel to each list
    fl.drain(NT);
    tl.drain(NT);
    if (p + 1 <= z1) {
      fetch(p + 2);
      fetch_words(p + 1);
    }
    fetch_orig(q + 1);
    if (do_a) {
      const double* Q0 = ring[(k + 1) % 5];
      const double* Q1 = ring[(k + 2) % 5];
      const double* Q2 = ring[(k + 3) % 5];
#pragma unroll 1
      for (int j = 0; j < 2; j++) {
        const int i = tid + NT * j;
        if (i < PN) {
          const int sx = i % kPX, sy = i / kPX, gx = x0 + sx - 1, gy = y0 + sy - 1;
          unsigned w = 0;
          if (p >= 0 && p < g.D && gx >= 0 && gx < g.W && gy >= 0 && gy < g.H)
            w = out_proposals<FAST>(s, Q0, Q1, Q2, (sx + 1) + kQX * (sy + 1), gx, gy, p,
                                    j ? ol1 : ol0, j ? ot1 : ot0);
          sring[k % 4][i] = w;
        }
      }
    }
    if (do_b && own_in) {
      const unsigned* S0 = sring[(k + 1) % 4];  // plane q - 1
      const unsigned* S1 = sring[(k + 2) % 4];  // plane q
      const unsigned* S2 = sring[(k + 3) % 4];  // plane q + 1
      const int sc = (lx + 1) + kPX * (ly + 1);
      const unsigned own = S1[sc];
      // bit 0: (orig - xi) - decp, bit 15: the lower midpoint; a neighbour in
      // direction d sends to this voxel along direction d ^ 1
      unsigned kinds = (own >> 14) & 0x8001u;
#pragma unroll
      for (int d = 0; d < kNbr; d++) {
        const unsigned* S = kSZ[d] < 0 ? S0 : (kSZ[d] > 0 ? S2 : S1);
        kinds |= (S[sc + kSX[d] + kPX * kSY[d]] >> (d ^ 1)) & 0x8001u;
      }
      const int t = x + g.W * (y + g.H * q);
      if (own >> 30) fl.push(t);
      if (kinds) {
        const double dv = ring[k % 5][(lx + 2) + kQX * (ly + 2)];
        long long key = s.sent;
        if (kinds & 1u) {
          const long long ka = okey(star_fix(s, og, dv));
          key = ka > key ? ka : key;
        }
        if (kinds & 0x8000u) {
          const long long kb =
              okey((s.step_q || s.f32) ? s.step : ((og - s.xi) + dv) / 2.0 - dv);
          key = kb > key ? kb : key;
        }
        if (key > s.sent) {
          s.delta[t] = key;
          tl.push(t);
        }
      }
    }
  }
  fl.flush();
  tl.flush();
}
template __global__ void k_scan_propose<true, kScanTY>(St, int*, int, int, int);
template __global__ void k_scan_propose<false, kScanTY>(St, int*, int, int, int);

// Follows every pointer out of a tile to its root (D7: every vertex is
// compressed).  A label without the top bit (kExit) is a root: found inside
// the tile by k_labels_tile or written here, so voxels resolved in their tile
// are only read.  A label with it names the next vertex on the steepest path;
// other lanes' writes only ever replace it by that path's root.  The walk is
// bounded by N.  Both labels of a voxel are walked together, so their loads
// are in flight at once.
__global__ void __launch_bounds__(256) k_chase(int n, int* ds, int* as) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int a0 = ds[v], b0 = as[v];
  if (a0 >= 0 && b0 >= 0) return;  // both resolved inside the tile
  int ra = a0 & ~kExit, rb = b0 & ~kExit;
  bool da = a0 >= 0, db = b0 >= 0;
  for (int step = 0; step < n && !(da && db); step++) {
    const int na = da ? 0 : __hip_atomic_load(&ds[ra], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nb = db ? 0 : __hip_atomic_load(&as[rb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!da) {
      da = na >= 0;
      ra = da ? na : na & ~kExit;
    }
    if (!db) {
      db = nb >= 0;
      rb = db ? nb : nb & ~kExit;
    }
  }
  if (a0 < 0) ds[v] = ra;
  if (b0 < 0) as[v] = rb;
}

// ------------------------------------------------------- CP structures ----
__device__ __forceinline__ void cas_best(int* slot, int cand, const double* f, bool want_max) {
  int cur = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  double fc = f[cand];
  while (true) {
    if (cur != kNone) {
      double fcur = f[cur];
      bool better = want_max ? larger(fc, cand, fcur, cur) : smaller(fc, cand, fcur, cur);
      if (!better) return;
    }
    int prev = atomicCAS(slot, cur, cand);
    if (prev == cur) return;
    cur = prev;
  }
}

__global__ void k_list_fill(const int* list, int n, int* arr, int val) {
  GRID_STRIDE(i, n) arr[list[i]] = val;
}

// Wavefront-wide best (value, id) over one (id, value) per lane; every lane
// receives the winner (and its value in fc).
__device__ __forceinline__ int wave_best(int cand, double& fc, bool want_max) {
  for (int o = 32; o > 0; o >>= 1) {
    int ob = __shfl_xor(cand, o);
    double of = __shfl_xor(fc, o);
    if (ob != kNone &&
        (cand == kNone || (want_max ? larger(of, ob, fc, cand) : smaller(of, ob, fc, cand)))) {
      cand = ob;
      fc = of;
    }
  }
  return cand;
}

// Ordered int64 key of a value with -0.0 folded into +0.0, so that equal
// values (the reference's `abs(a - b) == 0`) get equal keys.
__device__ __forceinline__ long long okey_c(double d) { return okey(d + 0.0); }

// Block-wide best (value, id) of one candidate per thread; returns it on
// thread 0.  All threads of the block must call; s_best is one int per
// wavefront of the block, in the calling kernel's LDS.
__device__ __forceinline__ int block_best(int cand, double fc, const double* f, bool want_max,
                                          int* s_best) {
  cand = wave_best(cand, fc, want_max);
  int tid = threadIdx.x + blockDim.x * threadIdx.y;
  if ((tid & 63) == 0) s_best[tid >> 6] = cand;
  __syncthreads();
  int best = kNone;
  if (tid == 0) {
    double fb = 0.0;
    const int nw = (blockDim.x * blockDim.y) >> 6;  // s_best holds one entry per wavefront
    for (int w = 0; w < nw; w++) {
      int c = s_best[w];
      if (c == kNone) continue;
      double fv = f[c];
      if (best == kNone || (want_max ? larger(fv, c, fb, best) : smaller(fv, c, fb, best))) {
        best = c;
        fb = fv;
      }
    }
  }
  __syncthreads();
  return best;
}

// One pass over the saddle lists (one lane per saddle and side), replacing
// computeCP + findAscPaths / findDesPaths +
// computelargestSaddlesForMax / computesmallestSaddlesForMin
// (src/ExtreM.cpp:1597-1679, 1920-2040, 2309-2371) without sorting.  The
// upper / lower stars come from flow, the lower-star masks of field f.
//  2-saddles: upper neighbours, ascent roots L2; F2 = lowest (value, id)
//   root (triplet[0]), La2 = highest (triplet[44]-1); every listed maximum
//   ends with the highest (value, id) saddle reaching it.
//  1-saddles: lower neighbours, descent roots L1; F1 = highest root, La1 =
//   lowest; every listed minimum ends with the lowest saddle.
// Per-extremum results live at the extremum's list position (ext_id): slot
// pass A raises the position's ordered value key (atomics without return)
// and records a (position, saddle) pair; k_slot_ids settles the id among the
// saddles holding the best key.  A root outside the lists (stale types, D10)
// has no position and is skipped; the global extremum is excluded where the
// slots are read.  With dec != 0, also the candidate masks of
// get_wrong_split_neighbors / get_wrong_join_neighbors
// (src/ExtreM.cpp:2405-2450): bit d set when the d-th neighbour is in the
// decp upper (lower) star and its root differs from the original root of the
// same-rank original upper (lower) neighbour.  Stencil loops are unrolled and
// all loads of a stage are issued before any is used.
template <int SIDE>
__device__ __forceinline__ void saddle_side(const St& s, int dec, const double* f, const int* L,
                                            const int* Lo, int v, unsigned star, unsigned ostar,
                                            const int* s_off, int* F, int* La, int o,
                                            unsigned& cand, int (&U)[kNbr], int& f_end,
                                            int& l_end, bool& thr_changed) {
  const Geo& g = s.g;
  int R[kNbr], O[kNbr];
  thr_changed = false;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    const bool in = (star >> d) & 1;
    R[d] = in ? L[v + g.off[d]] : kNone;
    O[d] = (dec && in) ? Lo[v + g.off[d]] : kNone;
  }
  unsigned uniq = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    bool dup = R[d] == kNone;
#pragma unroll
    for (int e = 0; e < d; e++) dup |= R[e] == R[d];
    if (!dup) uniq |= 1u << d;
  }
  double V[kNbr];
#pragma unroll
  for (int d = 0; d < kNbr; d++) V[d] = ((uniq >> d) & 1) ? f[R[d]] : 0.0;
  if (SIDE == 2 && s.thr > 0.0) {
    // the persistence filter of findAscPaths (src/ExtreM_cuda.cu:1877-1881):
    // a maximum closer than thr to the saddle leaves its triplet; a decp side
    // whose kept set differs from the original's is a changed side
    const double fv = f[v], ov = s.orig[v];
    unsigned keep = 0, okeep = 0;
#pragma unroll
    for (int d = 0; d < kNbr; d++) {
      if (!((uniq >> d) & 1)) continue;
      if (fabs(V[d] - fv) >= s.thr) keep |= 1u << d;
      if (dec && fabs(s.orig[R[d]] - ov) >= s.thr) okeep |= 1u << d;
    }
    thr_changed = dec && keep != okeep;
    uniq = keep;
  }
#pragma unroll
  for (int d = 0; d < kNbr; d++)
    U[d] = ((uniq >> d) & 1) ? R[d] : kNone;  // the distinct roots kept
  int lo = kNone, hi = kNone;
  double flo = 0.0, fhi = 0.0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (!((uniq >> d) & 1)) continue;
    if (lo == kNone || smaller(V[d], R[d], flo, lo)) { lo = R[d]; flo = V[d]; }
    if (hi == kNone || larger(V[d], R[d], fhi, hi)) { hi = R[d]; fhi = V[d]; }
  }
  f_end = SIDE == 2 ? lo : hi;
  l_end = SIDE == 2 ? hi : lo;
  if (o >= 0) {  // o < 0: a lane past the end of its list
    F[o] = f_end;
    La[o] = l_end;
  }
  cand = 0;
  if (!dec) return;
  if (star == ostar) {
    // decp star == original star (always so after c_loops): the k-th decp
    // neighbour is the k-th original neighbour
#pragma unroll
    for (int d = 0; d < kNbr; d++)
      if (R[d] != kNone && R[d] != O[d]) cand |= 1u << d;
  } else {
    unsigned st = star, os = ostar;
    while (st) {
      int d = __ffs(st) - 1;
      st &= st - 1;
      int od = os ? __ffs(os) - 1 : -1;
      os &= os - 1;
      int oroot = od < 0 ? kNone : Lo[v + s_off[od]];
      if (L[v + s_off[d]] != oroot) cand |= 1u << d;
    }
  }
}

// Slot pass A for one side: raises (maxima, even positions) or lowers (minima,
// odd positions) each distinct root's best saddle key (atomics whose result
// is not used, so none waits for its return) and records every
// (position, saddle) pair in the region of its 256-side chunk (cpairs,
// kChunkPairs entries, count in the LDS counter s_np: one LDS atomic per
// wavefront); k_slot_ids keeps the pairs that hold the final key.  Called by
// every lane of the block.
template <int SIDE>
__device__ __forceinline__ void emit_side(const St& s, int v, long long kv, int (&E)[kNbr],
                                          long long* key_slot, int2* cpairs, int* s_np) {
  int cnt = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (E[d] >= 0 && (E[d] & 1) == (SIDE == 1)) {
      if (SIDE == 2) atomicMax(&key_slot[E[d] >> 1], kv);
      else atomicMin(&key_slot[E[d] >> 1], kv);
      cnt++;
    } else {
      E[d] = -1;
    }
  }
  int pre = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(pre, o);
    if (__lane_id() >= o) pre += y;
  }
  const int total = __shfl(pre, 63);
  pre -= cnt;
  int base = 0;
  if (__lane_id() == 0 && total) base = atomicAdd(s_np, total);
  base = __shfl(base, 0) + pre;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (E[d] >= 0) {
      cpairs[base] = make_int2(E[d], v);
      base++;
    }
  }
}

// Slot index of a list position: maxima first, then minima.
__device__ __forceinline__ int slot_index(const St& s, int e) {
  return (e & 1) ? s.nmax + (e >> 1) : (e >> 1);
}

// Queues the slots whose original adjacent saddles include the saddle v
// (its original roots of this side: the extremum->saddle index of the
// original pass) for a new gather: v's gather key changed.
template <int SIDE>
__device__ __forceinline__ void regather_slots(const St& s, int v) {
  int x, y, z;
  coords(s.g, v, x, y, z);
  const unsigned key = c_cls_key[class_of(s.g, x, y, z)];
  const unsigned ol = s.olow[v];
  const unsigned ostar = SIDE == 2 ? key & ~ol : ol;
  const int* Lo = SIDE == 2 ? s.oDS : s.oAS;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (!((ostar >> d) & 1)) continue;
    const int e = s.ext_id[Lo[v + s.g.off[d]]];
    if (e < 0 || (e & 1) != (SIDE == 1)) continue;
    const int j = slot_index(s, e);
    if (atomicExch(&s.slstamp[j], s.iepoch) != s.iepoch) {
      // one counter atomic per wavefront for the lanes appending here
      const unsigned long long m = __ballot(1);
      const int lane = __lane_id(), leader = __ffsll((long long)m) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(&s.ctr[C_DSLOT], __popcll(m));
      base = __shfl(base, leader);
      s.dslot[base + __popcll(m & ((1ull << lane) - 1ull))] = j;
    }
  }
}

// A block's table of (slot, count) for the original pass's adjacency counts
// (neighbouring saddles share their roots): a slot's count is added here,
// or, when its probe finds no room, straight to the global count.
constexpr int kSlotCountN = 1024;
__device__ __forceinline__ void count_slot(int2* hc, int* gcount, int slot) {
  unsigned h = ((unsigned)slot * 2654435761u) >> 22;  // kSlotCountN = 2^10
  for (int p = 0; p < 16; p++) {
    const int old = atomicCAS(&hc[h].x, -1, slot);
    if (old == -1 || old == slot) {
      atomicAdd(&hc[h].y, 1);
      return;
    }
    h = (h + 1) & (kSlotCountN - 1);
  }
  atomicAdd(&gcount[slot], 1);
}

// One side for one saddle: the list entry i of S2 (SIDE 2) or S1 (SIDE 1).
// In kModeDec the wrong-neighbour candidates go to acl (act0) and the changed
// sides to cl (cln2 / cln1, for the next incremental pass); with inc, a side
// whose gather key changes queues its slots for a new gather.
template <int SIDE, class LA, class LC>
__device__ __forceinline__ void saddle_entry(const St& s, int mode, const double* f, const int* L,
                                             const uint16_t* flow, int* F, int* La, int i,
                                             bool act, const int* s_off, int2* cpairs,
                                             int* s_np, long long* key_slot, int inc, LA& acl,
                                             LC& cl, LC& wl, int2* hcount = nullptr) {
  const Geo& g = s.g;
  const int dec = mode == kModeDec;
  int v = 0;
  unsigned star = 0, ostar = 0;
  double fs = 0.0;
  // what a decp pass compares at the end, loaded with the list entry
  long long pvi = 0;
  int ofi = 0, oli = 0;
  uint8_t wfi = 1;
  if (act) {
    v = (SIDE == 2 ? s.S2 : s.S1)[i];
    if (dec) {
      pvi = (SIDE == 2 ? s.pv2 : s.pv1)[i];
      ofi = (SIDE == 2 ? s.of2 : s.of1)[i];
      oli = (SIDE == 2 ? s.ol2 : s.ol1)[i];
      wfi = s.wflag[SIDE == 2 ? i : s.n2 + i];
    }
    int x, y, z;
    coords(g, v, x, y, z);
    const unsigned key = c_cls_key[class_of(g, x, y, z)];
    const unsigned fl = flow[v], ol = s.olow[v];
    star = SIDE == 2 ? key & ~fl : key & fl;
    ostar = SIDE == 2 ? key & ~ol : ol;
    fs = f[v];
  }
  int E[kNbr];
  unsigned cand = 0;
  int f_end, l_end;
  bool thr_changed;
  saddle_side<SIDE>(s, dec, f, L, SIDE == 2 ? s.oDS : s.oAS, v, star, ostar, s_off, F, La,
                    act ? i : -1, cand, E, f_end, l_end, thr_changed);
  // list positions of the distinct roots (maxima even, minima odd; roots
  // outside the lists have none) where this saddle pushes or is indexed
  const bool changed = dec && (star != ostar || cand != 0 || thr_changed);
  if (!dec || changed) {
#pragma unroll
    for (int d = 0; d < kNbr; d++) {
      const int e = E[d] >= 0 ? s.ext_id[E[d]] : -1;
      E[d] = (e >= 0 && (e & 1) == (SIDE == 1)) ? e : -1;
    }
  } else {
#pragma unroll
    for (int d = 0; d < kNbr; d++) E[d] = -1;
  }
  if (mode == kModeOrig) {
    // count the (extremum, saddle) adjacencies and record them in the chunk's
    // region for k_csr_fill; every original saddle is gathered by its extrema
    int cnt = 0;
#pragma unroll
    for (int d = 0; d < kNbr; d++)
      if (E[d] >= 0) {
        E[d] = slot_index(s, E[d]);
        if (hcount) count_slot(hcount, s.csr_off, E[d]);
        else atomicAdd(&s.csr_off[E[d]], 1);
        cnt++;
      }
    int b = wave_reserve(s_np, cnt);
#pragma unroll
    for (int d = 0; d < kNbr; d++)
      if (E[d] >= 0) {
        cpairs[b++] = make_int2(E[d], i);
        E[d] = -1;
      }
    if (act) (SIDE == 2 ? s.pv2 : s.pv1)[i] = okey_c(fs);
  } else {
    // same star and the same root at every star neighbour: the root set is
    // the original one, gathered per extremum instead of pushed (its key at
    // its list position; a changed saddle leaves the empty key there)
    if (act) {
      long long* pv = SIDE == 2 ? s.pv2 : s.pv1;
      const long long nk =
          changed ? (SIDE == 2 ? (long long)0x8000000000000000ULL : 0x7fffffffffffffffLL)
                  : okey_c(fs);
      if (inc && pvi != nk) regather_slots<SIDE>(s, v);
      pv[i] = nk;
    }
  }
  emit_side<SIDE>(s, v, okey_c(fs), E, key_slot, cpairs, s_np);
  if (dec) {
    if (cand != 0) acl.push(make_int2(2 * v + (SIDE == 1), (int)cand));
    if (act && changed) cl.push(i);
    if (act) {  // the sides k_wrong_index has to look at
      const bool wm = f_end != ofi || l_end != oli;
      const int q = SIDE == 2 ? i : s.n2 + i;
      if (wm && !wfi) {
        s.wflag[q] = 1;
        wl.push(i);
      }
    }
  }
}

// The lists in chunks of 256 sides: chunks [0, nb2) are S2 entries, the rest
// S1 entries.  Each block takes a contiguous range of chunks, and the block
// order is remapped so each XCD walks one contiguous range (list order is
// spatial order, so the neighbour rows of a range stay in that XCD's L2).
// Triplet ends F2 / La2 / F1 / La1 are written at the list position, the
// slot pairs of chunk c (in kModeOrig its (slot, position) adjacencies) to
// its own region (count pcnt[c]), so no two blocks share a counter for them.
template <int MODE>
__global__ void __launch_bounds__(256, 4) k_saddle_list(St s, const double* f, const int* L2,
                                                     const int* L1, const uint16_t* flow,
                                                     int* F2, int* La2, int* F1, int* La1,
                                                     int2* pairs, int* pcnt, long long* kmax,
                                                     long long* kmin, int nb2, int nb) {
  __shared__ int s_off[kNbr];
  __shared__ int s_np;
  // the original pass counts its adjacencies per slot in the block first
  __shared__ int2 hc[MODE == kModeOrig ? kSlotCountN : 1];
  STAGED_LIST(sa, int2, 1024, s.act0, &s.ctr[C_ACT0]);
  STAGED_LIST(sc2, int, 1024, s.cln2, &s.ctr[C_CLN2]);
  STAGED_LIST(sc1, int, 1024, s.cln1, &s.ctr[C_CLN1]);
  STAGED_LIST(sw2, int, 1024, s.wl2, &s.ctr[C_WL2]);
  STAGED_LIST(sw1, int, 1024, s.wl1, &s.ctr[C_WL1]);
  const int tid = threadIdx.x;
  if (tid < kNbr) s_off[tid] = s.g.off[tid];
  if (MODE == kModeOrig)
    for (int j = tid; j < kSlotCountN; j += 256) hc[j] = make_int2(-1, 0);
  if (blockIdx.x == 0 && tid == 0) s.ctr[C_PCHUNK] = nb;
  // balanced ranges: every block (so every XCD) gets its share
  const long long L = xcd_remap(blockIdx.x, gridDim.x);
  const int c0 = (int)(L * nb / gridDim.x), c1 = (int)((L + 1) * nb / gridDim.x);
  for (int c = c0; c < c1; c++) {
    if (tid == 0) s_np = 0;
    __syncthreads();
    int2* cp = pairs + (long long)c * kChunkPairs;
    if (c < nb2) {
      const int i = c * 256 + tid;
      saddle_entry<2>(s, MODE, f, L2, flow, F2, La2, i, i < s.n2, s_off, cp, &s_np, kmax, 0, sa,
                      sc2, sw2, MODE == kModeOrig ? hc : nullptr);
    } else {
      const int i = (c - nb2) * 256 + tid;
      saddle_entry<1>(s, MODE, f, L1, flow, F1, La1, i, i < s.n1, s_off, cp, &s_np, kmin, 0, sa,
                      sc1, sw1, MODE == kModeOrig ? hc : nullptr);
    }
    __syncthreads();
    if (tid == 0) pcnt[c] = s_np;
    // (a side pushes at most one entry to each list)
    sa.drain(256);
    sc2.drain(256);
    sc1.drain(256);
    sw2.drain(256);
    sw1.drain(256);
  }
  if (MODE == kModeOrig) {
    __syncthreads();
    for (int j = tid; j < kSlotCountN; j += 256) {
      const int2 e = hc[j];
      if (e.x >= 0) atomicAdd(&s.csr_off[e.x], e.y);
    }
  }
  sa.flush();
  sc2.flush();
  sc1.flush();
  sw2.flush();
  sw1.flush();
}
template __global__ void k_saddle_list<kModeOrig>(St, const double*, const int*, const int*,
                                                  const uint16_t*, int*, int*, int*, int*, int2*,
                                                  int*, long long*, long long*, int, int);
template __global__ void k_saddle_list<kModeDec>(St, const double*, const int*, const int*,
                                                 const uint16_t*, int*, int*, int*, int*, int2*,
                                                 int*, long long*, long long*, int, int);

// The original extremum->saddle index from the adjacencies the kModeOrig
// pass recorded per chunk (one wavefront per chunk; the order within a slot's
// range is free: every reader takes a best over it).
// A block per chunk: each pair takes its rank among the chunk's pairs of its
// slot from an LDS table of (slot, count); then one global atomic per
// distinct slot of the chunk claims the slot's entries, and each pair is
// written at that base plus its rank.  A pair whose probe finds no room takes
// a global atomic of its own.
constexpr int kFillN = 1024;
// The pairs of the block's next chunk are loaded while the current one is
// placed (before its atomics and stores: a wait for a load also waits for the
// stores issued before it), and kept in registers through the placement.
__global__ void __launch_bounds__(256) k_csr_fill(St s, const int2* pairs, const int* pcnt,
                                                  int nb) {
  __shared__ int2 hc[kFillN];  // (slot, count), then (slot, base)
  constexpr int kPer = kChunkPairs / 256;
  const int tid = threadIdx.x;
  int2 nxt[kPer];
  int nn = 0;
  auto load = [&](int cc) {  // block-uniform
    nn = cc < nb ? pcnt[cc] : 0;
    const int2* P = pairs + (long long)cc * kChunkPairs;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int j = tid + 256 * k;
      nxt[k] = j < nn ? P[j] : make_int2(-1, 0);
    }
  };
  load(blockIdx.x);
  for (int c = blockIdx.x; c < nb; c += gridDim.x) {  // block-uniform
    const int n = nn;
    int2 pr[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) pr[k] = nxt[k];
    for (int j = tid; j < kFillN; j += 256) hc[j] = make_int2(-1, 0);
    __syncthreads();
    int hs[kPer], rk[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int j = tid + 256 * k;
      hs[k] = -1;
      rk[k] = 0;
      if (j >= n) continue;
      const int slot = pr[k].x;
      unsigned h = ((unsigned)slot * 2654435761u) >> 22;  // kFillN = 2^10
      for (int p = 0; p < 32; p++) {
        const int old = atomicCAS(&hc[h].x, -1, slot);
        if (old == -1 || old == slot) {
          hs[k] = (int)h;
          rk[k] = atomicAdd(&hc[h].y, 1);
          break;
        }
        h = (h + 1) & (kFillN - 1);
      }
    }
    __syncthreads();
    load(c + gridDim.x);
    for (int j = tid; j < kFillN; j += 256) {
      const int2 e = hc[j];
      if (e.x >= 0) hc[j].y = atomicAdd(&s.csr_cur[e.x], e.y);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const int j = tid + 256 * k;
      if (j >= n) continue;
      const int at = hs[k] >= 0 ? hc[hs[k]].y + rk[k] : atomicAdd(&s.csr_cur[pr[k].x], 1);
      s.csr[at] = pr[k].y;
    }
    __syncthreads();
  }
}

// Slot gather for the unchanged saddles of a decp pass: for slot index j
// (maxima, then minima) the best (key, vertex id) among the extremum's
// original adjacent saddles, over the keys pv2 / pv1 that the saddle pass
// left at the list positions ((key, id) is the (value, id) order).  pkey =
// the best key (the empty key if none), pid = its saddle vertex.  Slots with at
// most kPullShort adjacencies take a group of kPullGroup lanes; longer ones
// are cut into chunks of kPullChunk (k_long_slots), one block per chunk
// (k_slot_pull_long), joined by k_slot_pull_join.
constexpr int kPullShort = 256;
constexpr int kPullGroup = 4;
constexpr int kPullChunk = 1024;

// The candidates are list positions (of S2 for a maximum's slot, S1 for a
// minimum's): (key, saddle id) order with the ids looked up only where the
// keys are equal; the winner's id is looked up once, when it is stored.
__device__ __forceinline__ bool pull_better(const St& s, bool mx, long long ka, int pa,
                                            long long kb, int pb) {
  if (ka != kb) return mx ? ka > kb : ka < kb;
  const int* S = mx ? s.S2 : s.S1;
  const int ia = S[pa], ib = S[pb];
  return mx ? ia > ib : ia < ib;
}

__device__ __forceinline__ void pull_range(const St& s, bool mx, int k0, int k1, int step,
                                           long long& bk, int& bp) {
  const long long* pv = mx ? s.pv2 : s.pv1;
  int k = k0;
  for (; k + 3 * step < k1; k += 4 * step) {
    int p[4];
    long long q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) p[u] = s.csr[k + u * step];
#pragma unroll
    for (int u = 0; u < 4; u++) q[u] = pv[p[u]];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (bp < 0 || pull_better(s, mx, q[u], p[u], bk, bp)) { bk = q[u]; bp = p[u]; }
  }
  for (; k < k1; k += step) {
    const int p = s.csr[k];
    const long long q = pv[p];
    if (bp < 0 || pull_better(s, mx, q, p, bk, bp)) { bk = q; bp = p; }
  }
}

__device__ __forceinline__ void pull_group(const St& s, bool mx, int width, long long& bk,
                                           int& bp) {
  for (int o = 1; o < width; o <<= 1) {
    const long long ok = __shfl_xor(bk, o);
    const int op = __shfl_xor(bp, o);
    if (op >= 0 && (bp < 0 || pull_better(s, mx, ok, op, bk, bp))) { bk = ok; bp = op; }
  }
}

__device__ __forceinline__ void pull_store(const St& s, int j, bool mx, long long bk, int bp,
                                           long long* pkey, int* pid) {
  const long long empty = mx ? (long long)0x8000000000000000ULL : 0x7fffffffffffffffLL;
  const bool none = bp < 0 || bk == empty;
  pkey[j] = none ? empty : bk;
  pid[j] = none ? kNone : (mx ? s.S2 : s.S1)[bp];
}

__global__ void __launch_bounds__(256) k_slot_pull(St s, long long* pkey, int* pid) {
  const int n = s.nmax + s.nmin;
  const int gl = threadIdx.x & (kPullGroup - 1);
  const int groups = gridDim.x * (blockDim.x / kPullGroup);
  for (int j0 = blockIdx.x * (blockDim.x / kPullGroup); j0 < n; j0 += groups) {
    const int j = j0 + threadIdx.x / kPullGroup;
    int k0 = 0, k1 = 0;
    if (j < n) {
      k0 = s.csr_off[j];
      k1 = s.csr_off[j + 1];
    }
    const bool mine = k1 - k0 <= kPullShort;  // else taken by the chunked path
    const bool mx = j < s.nmax;
    long long bk = 0;
    int bp = -1;
    if (mine) pull_range(s, mx, k0 + gl, k1, kPullGroup, bk, bp);
    pull_group(s, mx, kPullGroup, bk, bp);
    if (j < n && mine && gl == 0) pull_store(s, j, mx, bk, bp, pkey, pid);
  }
}

// One block per chunk of a long slot: the chunk's best (key, position).
__global__ void __launch_bounds__(256) k_slot_pull_long(St s, long long* ckey, int* cpos) {
  __shared__ long long wk[4];
  __shared__ int wp[4];
  const int n = s.ctr[C_NCHUNK];
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    const int4 ch = s.chunks[c];  // {slot, k0, k1, -}
    const bool mx = ch.x < s.nmax;
    long long bk = 0;
    int bp = -1;
    pull_range(s, mx, ch.y + threadIdx.x, ch.z, 256, bk, bp);
    pull_group(s, mx, 64, bk, bp);
    if (__lane_id() == 0) {
      wk[threadIdx.x >> 6] = bk;
      wp[threadIdx.x >> 6] = bp;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < 4; w++)
        if (wp[w] >= 0 && (bp < 0 || pull_better(s, mx, wk[w], wp[w], bk, bp))) {
          bk = wk[w];
          bp = wp[w];
        }
      ckey[c] = bk;
      cpos[c] = bp;
    }
    __syncthreads();
  }
}

// Joins the chunks of every long slot.
__global__ void k_slot_pull_join(St s, const long long* ckey, const int* cpos, long long* pkey,
                                 int* pid) {
  const int n = s.ctr[C_NLONG];
  GRID_STRIDE(q, n) {
    const int2 l = s.longs[q];  // {slot, first chunk}
    const bool mx = l.x < s.nmax;
    const int nch = (s.csr_off[l.x + 1] - s.csr_off[l.x] + kPullChunk - 1) / kPullChunk;
    long long bk = 0;
    int bp = -1;
    for (int c = l.y; c < l.y + nch; c++)
      if (cpos[c] >= 0 && (bp < 0 || pull_better(s, mx, ckey[c], cpos[c], bk, bp))) {
        bk = ckey[c];
        bp = cpos[c];
      }
    pull_store(s, l.x, mx, bk, bp, pkey, pid);
  }
}

// Lists the slots with more than kPullShort adjacencies and cuts them into
// chunks of kPullChunk.
__global__ void k_long_slots(St s) {
  const int n = s.nmax + s.nmin;
  GRID_STRIDE(j, n) {
    const int k0 = s.csr_off[j], k1 = s.csr_off[j + 1];
    if (k1 - k0 <= kPullShort) continue;
    const int nch = (k1 - k0 + kPullChunk - 1) / kPullChunk;
    const int c0 = atomicAdd(&s.ctr[C_NCHUNK], nch);
    s.longs[atomicAdd(&s.ctr[C_NLONG], 1)] = make_int2(j, c0);
    for (int c = 0; c < nch; c++)
      s.chunks[c0 + c] = make_int4(j, k0 + c * kPullChunk, min(k1, k0 + (c + 1) * kPullChunk), 0);
  }
}

// Final slot keys: the better of the gathered and the pushed key; the slot
// starts with the gathered saddle when it holds the final key (k_slot_ids
// then adds the pushed saddles holding it).
// Slot index j (maxima, then minima): the pushed key merged with the
// gathered one; the slot is the gathered saddle when its key stands (ties
// among pushed saddles follow in k_slot_ids).  Idempotent, so a slot may be
// merged more than once.
__device__ __forceinline__ void merge_slot(const St& s, int j, long long* kmax, long long* kmin,
                                           int* slot_max, int* slot_min, const long long* pkey,
                                           const int* pid) {
  if (j < s.nmax) {
    const long long k = kmax[j] > pkey[j] ? kmax[j] : pkey[j];
    kmax[j] = k;
    slot_max[j] = (pid[j] != kNone && pkey[j] == k) ? pid[j] : kNone;
  } else {
    const int i = j - s.nmax;
    const long long k = kmin[i] < pkey[j] ? kmin[i] : pkey[j];
    kmin[i] = k;
    slot_min[i] = (pid[j] != kNone && pkey[j] == k) ? pid[j] : 0x7fffffff;
  }
}

__global__ void k_slot_merge(St s, long long* kmax, long long* kmin, int* slot_max,
                             int* slot_min, const long long* pkey, const int* pid) {
  GRID_STRIDE(j, s.nmax + s.nmin) merge_slot(s, j, kmax, kmin, slot_max, slot_min, pkey, pid);
}

// In-place exclusive scan of a[0..n) by one block of 1024 threads; a[n]
// receives the total.
__global__ void __launch_bounds__(1024) k_scan_single(int* a, int n) {
  __shared__ int sh[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b0 = min(n, t * per), b1 = min(n, b0 + per);
  int sum = 0;
  for (int b = b0; b < b1; b++) sum += a[b];
  sh[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int x = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int run = sh[t] - sum;
  for (int b = b0; b < b1; b++) {
    int x = a[b];
    a[b] = run;
    run += x;
  }
  if (t == 1023) a[n] = sh[1023];
}

// Best (value, id) of two vertex lists under f: out[block] = the highest of
// inmax's grid-stride share and the lowest of inmin's (kNone entries skipped).
// Run over MX / MN, then over the per-block results with one block.
__global__ void __launch_bounds__(256) k_ext_reduce(const int* inmax, int nmax, const int* inmin,
                                                    int nmin, const double* f, int* outmax,
                                                    int* outmin) {
  int cmax = kNone, cmin = kNone;
  double fmax = 0, fmin = 0;
  GRID_STRIDE(i, nmax) {
    int a = inmax[i];
    if (a != kNone) {
      double fa = f[a];
      if (cmax == kNone || larger(fa, a, fmax, cmax)) { cmax = a; fmax = fa; }
    }
  }
  GRID_STRIDE(i, nmin) {
    int b = inmin[i];
    if (b != kNone) {
      double fb = f[b];
      if (cmin == kNone || smaller(fb, b, fmin, cmin)) { cmin = b; fmin = fb; }
    }
  }
  __shared__ int s_best[4];
  int gm = block_best(cmax, fmax, f, true, s_best);
  int gn = block_best(cmin, fmin, f, false, s_best);
  if (threadIdx.x == 0) {
    outmax[blockIdx.x] = gm;
    outmin[blockIdx.x] = gn;
  }
}

// Slot pass B: among the recorded (position, saddle) pairs, the saddles
// whose value equals the position's best value key settle the id (highest id
// for a maximum's slot, lowest for a minimum's), giving the (value, id) order
// of computelargestSaddlesForMax / computesmallestSaddlesForMin.  One
// wavefront per chunk region of the last saddle launch.
__device__ __forceinline__ void slot_ids_body(const St& s, const int2* pairs, const int* pcnt,
                                              const double* f, const long long* kmax,
                                              const long long* kmin, int* slot_max, int* slot_min) {
  const int nc = s.ctr[C_PCHUNK];
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nc; c += nw) {
    const int n = pcnt[c];
    const int2* P = pairs + (long long)c * kChunkPairs;
    for (int j = lane; j < n; j += 64) {
      const int2 pr = P[j];
      const int e = pr.x >> 1;
      const long long k = okey_c(f[pr.y]);
      if (pr.x & 1) {
        if (k == kmin[e]) atomicMin(&slot_min[e], pr.y);
      } else {
        if (k == kmax[e]) atomicMax(&slot_max[e], pr.y);
      }
    }
  }
}
__global__ void __launch_bounds__(256) k_slot_ids(St s, const int2* pairs, const int* pcnt,
                                                  const double* f, const long long* kmax,
                                                  const long long* kmin, int* slot_max,
                                                  int* slot_min) {
  slot_ids_body(s, pairs, pcnt, f, kmax, kmin, slot_max, slot_min);
}

// ext_id[v] = 2 * (position in MX) for maxima, 2 * (position in MN) + 1 for
// minima (set once per run; -1 elsewhere).
__global__ void k_ext_index(St s) {
  GRID_STRIDE(i, s.nmax) s.ext_id[s.MX[i]] = 2 * i;
  GRID_STRIDE(i, s.nmin) s.ext_id[s.MN[i]] = 2 * i + 1;
}

// Empty slot keys and slots for a saddle pass, and its list counters (read
// by the launches before it, appended to by the ones after).
__global__ void k_slot_init(St s, long long* kmax, int* smax, long long* kmin, int* smin) {
  GRID_STRIDE(i, s.nmax) {
    kmax[i] = (long long)0x8000000000000000ULL;
    smax[i] = kNone;
  }
  GRID_STRIDE(i, s.nmin) {
    kmin[i] = 0x7fffffffffffffffLL;
    smin[i] = 0x7fffffff;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.ctr[C_ACT0] = 0;
    s.ctr[C_CLN2] = 0;
    s.ctr[C_CLN1] = 0;
    s.ctr[C_DSLOT] = 0;
  }
}

__global__ void k_list_fill64(const int* list, int n, long long* arr, long long val) {
  GRID_STRIDE(i, n) arr[list[i]] = val;
}

// ------------------------------------------------ incremental CP pass ----
// A CP pass of decp after few edits recomputes only what the edits can
// reach, and leaves the state the full pass (label_pass + slot_pass) would:
//  1. k_inc_ptrs: lower-star masks and steepest directions of the changed
//     vertices and their neighbours, the only vertices whose star or
//     pointers can differ.
//  2. k_inc_taint + k_inc_walk: the new root of every vertex whose pointer
//     changed, by walking its new path up to a vertex whose old root is not
//     the old root of any changed pointer (no changed pointer lies on that
//     vertex's path, so its old root stands) or that is a changed pointer
//     whose new root is known.
//  3. k_inc_prop: every other vertex whose root changes has a path that
//     enters a changed pointer whose root changed, and takes that root;
//     these are found breadth-first over the pointer tree (the children of
//     u are the neighbours whose pointer names u), stopping at changed
//     pointers, which carry their own root.
//  4. k_inc_mark / k_inc_mark_csr: the saddle sides whose inputs changed --
//     their star (lower mask), their value, the root of a star neighbour, the
//     value of one of their roots -- and every side that differed from the
//     original in the last pass (slot pushes and wrong-neighbour candidates
//     are rebuilt each pass).
//  5. k_inc_saddles: those sides through saddle_entry; k_inc_pull: the slots
//     whose gathered keys changed.

// One lane per (changed vertex, candidate): candidate 0 is the vertex,
// 1 + d its neighbour in direction d; each vertex once per epoch.  New lower
// masks and steepest directions (the order of k_labels_tile); lists the
// vertices whose mask (al), up pointer (cu) or down pointer (cd) changed.
__device__ __forceinline__ void inc_ptrs_body(const St& s, int epoch) {
  STAGED_LIST(la, int, 512, s.al, &s.ctr[C_AL]);
  STAGED_LIST(lu, int, 512, s.cu, &s.ctr[C_CU]);
  STAGED_LIST(ld, int, 512, s.cd, &s.ctr[C_CD]);
  const long long n = (long long)s.ctr[C_LCHG] * (kNbr + 1);
  // (one entry per item at most to each list)
  block_rounds(n, 1, [&](long long k) {
    const int i = (int)(k / (kNbr + 1)), j = (int)(k - (long long)i * (kNbr + 1));
    int u = s.lchg[i];
    int x, y, z;
    coords(s.g, u, x, y, z);
    int cls = class_of(s.g, x, y, z);
    if (j == 0) {  // the global extremum of decp changed value: k_ext_reduce runs
      if (u == s.ctr[C_DGMAX] || u == s.ctr[C_DGMIN]) s.ctr[C_EXTD] = 1;
    } else {
      if (!((c_cls_key[cls] >> (j - 1)) & 1)) return;
      u += s.g.off[j - 1];
      cls = class_of(s.g, x + c_dx[j - 1], y + c_dy[j - 1], z + c_dz[j - 1]);
    }
    if (atomicExch(&s.stamp[u], epoch) == epoch) return;
    const int key = c_cls_key[cls];
    const double c = s.decp[u];
    unsigned lm = 0;
    int up = u, dn = u, du = -1, dd = -1;
    double fu = c, fd = c;
#pragma unroll
    for (int d = 0; d < kNbr; d++) {
      if (!((key >> d) & 1)) continue;
      const int r = u + s.g.off[d];
      const double fr = s.decp[r];
      if (fr < c || (fr == c && r < u)) lm |= 1u << d;
      if (fu < fr || (fu == fr && up < r)) { up = r; fu = fr; du = d; }
      if (fd > fr || (fd == fr && dn > r)) { dn = r; fd = fr; dd = d; }
    }
    const unsigned nd = (unsigned)(du + 1) | (unsigned)(dd + 1) << 4;
    const unsigned od = s.pdir[u];
    if (lm != s.dlow[u]) {
      s.dlow[u] = (uint16_t)lm;
      la.push(u);
    }
    unsigned cf = 0;
    if ((nd ^ od) & 15u) {
      cf |= 1;
      lu.push(u);
    }
    if ((nd ^ od) >> 4) {
      cf |= 2;
      ld.push(u);
    }
    if (cf) {
      s.pdir[u] = (uint8_t)nd;
      s.cflag[u] = (uint8_t)cf;
    }
  }, la, lu, ld);
  la.flush();
  lu.flush();
  ld.flush();
}
__global__ void __launch_bounds__(256) k_inc_ptrs(St s, int epoch) {
  inc_ptrs_body(s, epoch);
}

// The old roots of the changed pointers (kept in ou / od, flagged in tu / td).
// Also: a listed extremum other than the current global one changed value,
// so the global maximum / minimum of decp is the better of it and the current
// one (when the current one changed, k_inc_ptrs set C_EXTD, both are cleared
// here and k_inc_mark recomputes them).
__device__ __forceinline__ void inc_taint_body(const St& s) {
  if (s.ctr[C_EXTD]) {  // k_inc_mark recomputes them from scratch
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      s.ctr[C_DGMAX] = kNone;
      s.ctr[C_DGMIN] = kNone;
    }
  } else {
    const int n = s.ctr[C_LCHG];
    GRID_STRIDE(w, n) {
      const int u = s.lchg[w];
      const int e = s.ext_id[u];
      if (e < 0) continue;
      if ((e & 1) == 0) cas_best(&s.ctr[C_DGMAX], u, s.decp, true);
      else cas_best(&s.ctr[C_DGMIN], u, s.decp, false);
    }
  }
  const int nu = s.ctr[C_CU], nd = s.ctr[C_CD];
  GRID_STRIDE(k, nu) {
    const int r = s.dDS[s.cu[k]];
    s.ou[k] = r;
    s.tu[r] = 1;
  }
  GRID_STRIDE(k, nd) {
    const int r = s.dAS[s.cd[k]];
    s.od[k] = r;
    s.td[r] = 1;
  }
}
__global__ void k_inc_taint(St s) {
  inc_taint_body(s);
}

// New root of each changed pointer, written to its root array; the pointers
// whose root changed start the propagation queues (qu / qd).  A path strictly
// increases (decreases) in (value, id), so it ends within N steps.  A root
// array entry read on the way is either the old root (tainted when the
// vertex's path holds a changed pointer, so the walk goes on) or the final new
// root of a changed pointer, so either answer it gives is exact.
__device__ __forceinline__ void inc_walk_body(const St& s) {
  STAGED_LIST(su, int, 512, s.qu, &s.ctr[C_QU]);
  STAGED_LIST(sd, int, 512, s.qd, &s.ctr[C_QD]);
  const int nu = s.ctr[C_CU], n = nu + s.ctr[C_CD];
  block_rounds(n, 1, [&](long long kk) {
    const int k = (int)kk;
    const bool up = k < nu;
    const int c = up ? s.cu[k] : s.cd[k - nu];
    int* R = up ? s.dDS : s.dAS;
    const uint8_t* T = up ? s.tu : s.td;
    const int sh = up ? 0 : 4;
    int a = c, root = -1;
    for (int step = 0; step < s.g.N && root < 0; step++) {
      const int nib = (s.pdir[a] >> sh) & 15;
      if (nib == 0) {
        root = a;
        break;
      }
      a += s.g.off[nib - 1];
      const int ra = R[a];
      if (!T[ra]) root = ra;
    }
    if (root < 0) {
      s.ctr[C_IBAIL] = 1;
      root = a;
    }
    R[c] = root;
    if (root != (up ? s.ou[k] : s.od[k - nu])) {
      if (up) su.push(c);
      else sd.push(c);
    }
  }, su, sd);
  su.flush();
  sd.flush();
  // the last block to finish opens the first propagation level
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&s.ctr[C_SWEEP], 1) == (int)gridDim.x - 1) {
      __threadfence();
      s.ctr[C_QHU] = __hip_atomic_load(&s.ctr[C_QU], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.ctr[C_QHD] = __hip_atomic_load(&s.ctr[C_QD], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.ctr[C_SWEEP] = 0;
    }
  }
}
__global__ void __launch_bounds__(256) k_inc_walk(St s) {
  inc_walk_body(s);
}

// The children of u (neighbours whose pointer of this side names u) that are
// not changed pointers themselves.
__device__ __forceinline__ unsigned prop_children(const St& s, int u, int sh, unsigned cbit) {
  int x, y, z;
  coords(s.g, u, x, y, z);
  const int key = c_cls_key[class_of(s.g, x, y, z)];
  unsigned pd[kNbr];
#pragma unroll
  for (int d = 0; d < kNbr; d++)  // all loads issued before any is used
    pd[d] = ((key >> d) & 1) ? s.pdir[u + s.g.off[d]] : 0u;
  unsigned kids = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++)
    if ((int)((pd[d] >> sh) & 15) == (d ^ 1) + 1) kids |= 1u << d;
  unsigned cf[kNbr];
#pragma unroll
  for (int d = 0; d < kNbr; d++) cf[d] = ((kids >> d) & 1) ? s.cflag[u + s.g.off[d]] : 0u;
#pragma unroll
  for (int d = 0; d < kNbr; d++)
    if (cf[d] & cbit) kids &= ~(1u << d);
  return kids;
}

// The vertices below each changed pointer whose root changed take its new
// root.  Each queued vertex already holds its final root, so the queue may be
// worked in any order; it is worked level by level, a level being what the
// queue held when the previous one was done: first by whole-grid sweeps (one
// level each, the last block to finish opens the next level), then by
// k_inc_prop in one block for the deep remainder.  A vertex has one parent,
// so none is queued twice.
__global__ void __launch_bounds__(256) k_inc_sweep(St s) {
  const int lou = s.ctr[C_QLU], hiu = s.ctr[C_QHU], lod = s.ctr[C_QLD], hid = s.ctr[C_QHD];
  const int nu = hiu - lou, n = nu + (hid - lod);
  const int stride = gridDim.x * blockDim.x;
  for (int k0 = blockIdx.x * blockDim.x; k0 < n; k0 += stride) {  // uniform per block
    const int k = k0 + threadIdx.x;
    const bool up = k < nu;  // lanes of one wave may differ: appends go per side
    int u = 0, root = 0;
    unsigned kids = 0;
    if (k < n) {
      u = up ? s.qu[lou + k] : s.qd[lod + k - nu];
      root = (up ? s.dDS : s.dAS)[u];
      kids = prop_children(s, u, up ? 0 : 4, up ? 1u : 2u);
    }
    int* R = up ? s.dDS : s.dAS;
#pragma unroll
    for (int d = 0; d < kNbr; d++)
      if ((kids >> d) & 1) R[u + s.g.off[d]] = root;
    const int cu = up ? __popc(kids) : 0, cd = up ? 0 : __popc(kids);
    int bu = wave_reserve(&s.ctr[C_QU], cu);
    int bd = wave_reserve(&s.ctr[C_QD], cd);
    int* Q = up ? s.qu : s.qd;
    int b = up ? bu : bd;
#pragma unroll
    for (int d = 0; d < kNbr; d++)
      if ((kids >> d) & 1) {
        if (b < s.g.N) Q[b] = u + s.g.off[d];
        else s.ctr[C_IBAIL] = 1;
        b++;
      }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&s.ctr[C_SWEEP], 1) == (int)gridDim.x - 1) {  // the last block
      __threadfence();
      s.ctr[C_QLU] = hiu;
      s.ctr[C_QHU] = __hip_atomic_load(&s.ctr[C_QU], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.ctr[C_QLD] = hid;
      s.ctr[C_QHD] = __hip_atomic_load(&s.ctr[C_QD], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s.ctr[C_SWEEP] = 0;
    }
  }
}

// The remainder of the propagation in one block, level by level from
// [C_QLx, C_QHx) to an empty level.  Every lane of the block walks the same
// rounds (the queue appends are wavefront collectives).  qu / qd end up
// listing every vertex whose root changed.
__device__ __forceinline__ void inc_prop_body(const St& s, int qcap) {
  __shared__ int qn;
  const int tid = threadIdx.x;
  for (int side = 0; side < 2; side++) {
    const bool up = side == 0;
    int* R = up ? s.dDS : s.dAS;
    int* Q = up ? s.qu : s.qd;
    const int sh = up ? 0 : 4;
    const unsigned cbit = up ? 1u : 2u;
    if (tid == 0) qn = s.ctr[up ? C_QU : C_QD];
    __syncthreads();
    int lo = s.ctr[up ? C_QLU : C_QLD], hi = min(qn, qcap);
    __syncthreads();
    while (lo < hi) {
      for (int k0 = lo; k0 < hi; k0 += 1024) {
        const int k = k0 + tid;
        int u = 0, root = 0;
        unsigned kids = 0;
        if (k < hi) {
          u = Q[k];
          root = R[u];
          kids = prop_children(s, u, sh, cbit);
        }
        int b = wave_reserve(&qn, __popc(kids));
#pragma unroll
        for (int d = 0; d < kNbr; d++) {
          if (!((kids >> d) & 1)) continue;
          const int nb = u + s.g.off[d];
          R[nb] = root;
          if (b < qcap) Q[b] = nb;
          b++;
        }
      }
      __syncthreads();
      lo = hi;
      hi = min(qn, qcap);
      __syncthreads();
    }
    if (tid == 0) {
      s.ctr[up ? C_QU : C_QD] = min(qn, qcap);
      if (qn > qcap) s.ctr[C_IBAIL] = 1;
    }
    __syncthreads();
  }
}
__global__ void __launch_bounds__(1024) k_inc_prop(St s, int qcap) {
  inc_prop_body(s, qcap);
}

// Staged entries of the side-marking lists (an item of k_inc_mark marks up to
// 14 sides; a wavefront of k_inc_mark_csr up to 1024).
constexpr int kMarkStage = 4096;

// Queues side q2 of S2 (two) or S1 for the saddle recompute, once per pass.
template <class L>
__device__ __forceinline__ void mark_side(const St& s, int q2, bool two, L& l2, L& l1) {
  const int q = two ? q2 : s.n2 + q2;
  if (atomicExch(&s.sstamp[q], s.iepoch) == s.iepoch) return;
  if (two) l2.push(q2);
  else l1.push(q2);
}

// Sides of saddles whose star or value changed (al, lchg), of saddles next to
// a vertex whose root changed (qu: 2-saddles read up roots, qd: 1-saddles
// down roots), and the changed sides of the last pass (cl2, cl1).  Also the
// chunks of k_inc_mark_csr, empty slot keys and slots for this pass's pushes
// (as k_slot_init; no kernel before the pushes reads them), and the decp
// global extrema when k_inc_taint cleared them.
__device__ __forceinline__ void inc_mark_body(const St& s, long long* kmax, long long* kmin,
                                              int all_slots) {
  if (s.ctr[C_IBAIL]) return;
  // the global maximum / minimum of decp over the lists, when the current one
  // changed value (block bests into the cleared counters; the block barriers
  // are reached by every thread either way, over empty ranges otherwise)
  {
    const bool ext = s.ctr[C_EXTD] != 0;
    const int nmx = ext ? s.nmax : 0, nmn = ext ? s.nmin : 0;
    int cmax = kNone, cmin = kNone;
    double fmax = 0.0, fmin = 0.0;
    GRID_STRIDE(i, nmx) {
      const int a = s.MX[i];
      const double fa = s.decp[a];
      if (cmax == kNone || larger(fa, a, fmax, cmax)) { cmax = a; fmax = fa; }
    }
    GRID_STRIDE(i, nmn) {
      const int b = s.MN[i];
      const double fb = s.decp[b];
      if (cmin == kNone || smaller(fb, b, fmin, cmin)) { cmin = b; fmin = fb; }
    }
    __shared__ int s_best[16];  // blocks of up to 1024 threads
    const int gm = block_best(cmax, fmax, s.decp, true, s_best);
    const int gn = block_best(cmin, fmin, s.decp, false, s_best);
    if (threadIdx.x == 0) {
      if (gm != kNone) cas_best(&s.ctr[C_DGMAX], gm, s.decp, true);
      if (gn != kNone) cas_best(&s.ctr[C_DGMIN], gn, s.decp, false);
    }
  }

  STAGED_LIST(l2, int, kMarkStage, s.dsl2, &s.ctr[C_DSL2]);
  STAGED_LIST(l1, int, kMarkStage, s.dsl1, &s.ctr[C_DSL1]);
  const long long nal = s.ctr[C_AL], nch = s.ctr[C_LCHG], nqu = s.ctr[C_QU],
                  nqd = s.ctr[C_QD], ncl2 = s.ctr[C_CL2], ncl1 = s.ctr[C_CL1];
  const long long n = nal + nch + nqu + nqd + ncl2 + ncl1;
  // (an item marks up to 14 sides: the lists hold kMarkStage entries and are
  // emptied between rounds)
  block_rounds(n, kNbr, [&](long long k) {
    long long t = k;
    if (t < nal + nch) {
      const int u = t < nal ? s.al[t] : s.lchg[t - nal];
      const int p2 = s.spos2[u], p1 = s.spos1[u];
      if (p2 >= 0) mark_side(s, p2, true, l2, l1);
      if (p1 >= 0) mark_side(s, p1, false, l2, l1);
      return;
    }
    t -= nal + nch;
    if (t < nqu + nqd) {
      const bool up = t < nqu;
      const int u = up ? s.qu[t] : s.qd[t - nqu];
      const int* P = up ? s.spos2 : s.spos1;
      int x, y, z;
      coords(s.g, u, x, y, z);
      const int key = c_cls_key[class_of(s.g, x, y, z)];
      int p[kNbr];
#pragma unroll
      for (int d = 0; d < kNbr; d++) p[d] = ((key >> d) & 1) ? P[u + s.g.off[d]] : -1;
#pragma unroll
      for (int d = 0; d < kNbr; d++)
        if (p[d] >= 0) mark_side(s, p[d], up, l2, l1);
      return;
    }
    t -= nqu + nqd;
    if (t < ncl2) mark_side(s, s.cl2[t], true, l2, l1);
    else mark_side(s, s.cl1[t - ncl2], false, l2, l1);
  }, l2, l1);
  l2.flush();
  l1.flush();
  // the extremum->saddle index ranges of the changed extrema that are still
  // roots, in chunks of 1024 (uniform rounds: the chunk counts are summed per
  // wavefront)
  const int stride = gridDim.x * blockDim.x;
  for (int w0 = blockIdx.x * blockDim.x; w0 < nch; w0 += stride) {
    const int w = w0 + threadIdx.x;
    int k0 = 0, k1 = 0, mx = 0;
    if (w < nch) {
      const int r = s.lchg[w];
      const int e = s.ext_id[r];
      if (e >= 0) {
        mx = (e & 1) == 0;
        if ((mx ? s.dDS[r] : s.dAS[r]) == r) {
          const int j = slot_index(s, e);
          k0 = s.csr_off[j];
          k1 = s.csr_off[j + 1];
        }
      }
    }
    const int nck = k1 > k0 ? (k1 - k0 + 1023) >> 10 : 0;
    int pre = nck;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(pre, o);
      if (__lane_id() >= o) pre += y;
    }
    const int total = __shfl(pre, 63);
    if (total == 0) continue;
    int b = 0;
    if (__lane_id() == 63) b = atomicAdd(&s.ctr[C_NITEM], total);
    b = __shfl(b, 63) + pre - nck;
    for (int c = 0; c < nck; c++)
      s.citems[b + c] = make_int4(k0 + (c << 10), min(k1, k0 + ((c + 1) << 10)), mx, 0);
  }
  if (all_slots) {
    GRID_STRIDE(i, s.nmax) {
      kmax[i] = (long long)0x8000000000000000ULL;
      s.dlarg[i] = kNone;
    }
    GRID_STRIDE(i, s.nmin) {
      kmin[i] = 0x7fffffffffffffffLL;
      s.dsmall[i] = 0x7fffffff;
    }
  } else {
    // every other slot key is the merged one of the last pass with no push
    // in it, i.e. the gathered key, which a push of this pass may raise
    const int np = s.ctr[C_PPREV];
    GRID_STRIDE(i, np) {
      const int j = s.pprev[i];
      if (j < s.nmax) kmax[j] = (long long)0x8000000000000000ULL;
      else kmin[j - s.nmax] = 0x7fffffffffffffffLL;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    s.ctr[C_ACT0] = 0;
    s.ctr[C_CLN2] = 0;
    s.ctr[C_CLN1] = 0;
    s.ctr[C_DSLOT] = 0;
  }
}
__global__ void __launch_bounds__(256) k_inc_mark(St s, long long* kmax, long long* kmin,
                                                  int all_slots) {
  inc_mark_body(s, kmax, kmin, all_slots);
}

// Slot merge of an incremental pass: every slot (all_slots, the first pass
// after a full one), else the slots pushed by the last pass (reset by
// k_inc_mark) and the gathered-again ones; always the slots of this pass's
// pair records, which are listed (deduplicated) for the next pass.  Other
// slots keep last pass's keys and slots: nothing they merge changed.
__device__ __forceinline__ void inc_merge_body(const St& s, long long* kmax, long long* kmin,
                                               const long long* pkey, const int* pid,
                                               const int2* pairs, const int* pcnt, int all_slots) {
  STAGED_LIST(pl, int, 1024, s.pnext, &s.ctr[C_PNEXT]);
  if (all_slots) {
    GRID_STRIDE(j, s.nmax + s.nmin)
        merge_slot(s, j, kmax, kmin, s.dlarg, s.dsmall, pkey, pid);
  } else {
    const int np = s.ctr[C_PPREV], nd = s.ctr[C_DSLOT];
    GRID_STRIDE(i, np) merge_slot(s, s.pprev[i], kmax, kmin, s.dlarg, s.dsmall, pkey, pid);
    GRID_STRIDE(i, nd) merge_slot(s, s.dslot[i], kmax, kmin, s.dlarg, s.dsmall, pkey, pid);
  }
  const int nc = s.ctr[C_PCHUNK];
  const int lane = threadIdx.x & 63;
  const int nw = (gridDim.x * blockDim.x) >> 6;
  for (int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; c < nc; c += nw) {
    const int n = pcnt[c];
    const int2* P = pairs + (long long)c * kChunkPairs;
    for (int k = lane; k < n; k += 64) {
      const int2 pr = P[k];
      const int j = (pr.x & 1) ? s.nmax + (pr.x >> 1) : (pr.x >> 1);
      merge_slot(s, j, kmax, kmin, s.dlarg, s.dsmall, pkey, pid);
      if (atomicExch(&s.pstamp[j], s.iepoch) != s.iepoch) pl.push(j);
    }
  }
  pl.flush();
}
__global__ void __launch_bounds__(256) k_inc_merge(St s, long long* kmax, long long* kmin,
                                                   const long long* pkey, const int* pid,
                                                   const int2* pairs, const int* pcnt,
                                                   int all_slots) {
  inc_merge_body(s, kmax, kmin, pkey, pid, pairs, pcnt, all_slots);
}

// The saddles adjacent in the original to a changed vertex that is a listed
// extremum and still a root of decp compare its value, so their sides are
// queued; the other saddles with this root differ from the original and are
// queued as last pass's changed sides or by a root change next to them.
// k_inc_mark cuts those extrema's index ranges into chunks of 1024,
// k_inc_mark_csr queues one chunk per wavefront.
__device__ __forceinline__ void inc_mark_csr_body(const St& s) {
  if (s.ctr[C_IBAIL]) return;
  STAGED_LIST(l2, int, kMarkStage, s.dsl2, &s.ctr[C_DSL2]);
  STAGED_LIST(l1, int, kMarkStage, s.dsl1, &s.ctr[C_DSL1]);
  const int n = s.ctr[C_NITEM];
  const int lane = threadIdx.x & 63;
  const int bw = blockDim.x >> 6, nw = gridDim.x * bw;
  // block-uniform rounds of one chunk per wavefront (up to 1024 marks each),
  // the lists emptied between rounds
  for (int w0 = blockIdx.x * bw; w0 < n; w0 += nw) {
    const int w = w0 + (threadIdx.x >> 6);
    if (w < n) {
      const int4 it = s.citems[w];
      for (int k = it.x + lane; k < it.y; k += 64) mark_side(s, s.csr[k], it.z != 0, l2, l1);
    }
    __syncthreads();
    l2.drain(1024 * bw);
    l1.drain(1024 * bw);
  }
  l2.flush();
  l1.flush();
}
__global__ void __launch_bounds__(256) k_inc_mark_csr(St s) {
  inc_mark_csr_body(s);
}

// The queued sides through saddle_entry (kModeDec, with slot re-gather
// marking), in chunks of 256 queued sides (S2 sides first, each part rounded
// up to whole chunks), a contiguous range of chunks per block, slot pairs in
// the chunks' own regions as in k_saddle_list.  A block of NT threads works
// NT / 256 chunks of its range at a time, one per 256-thread group.
template <int NT>
__device__ __forceinline__ void inc_saddles_body(const St& s, const double* f, int2* pairs,
                                                 int* pcnt, long long* kmax, long long* kmin) {
  constexpr int kG = NT / 256;
  __shared__ int s_off[kNbr];
  __shared__ int s_np[kG];
  STAGED_LIST(sa, int2, 1024, s.act0, &s.ctr[C_ACT0]);
  STAGED_LIST(sc2, int, 1024, s.cln2, &s.ctr[C_CLN2]);
  STAGED_LIST(sc1, int, 1024, s.cln1, &s.ctr[C_CLN1]);
  STAGED_LIST(sw2, int, 1024, s.wl2, &s.ctr[C_WL2]);
  STAGED_LIST(sw1, int, 1024, s.wl1, &s.ctr[C_WL1]);
  const int tid = threadIdx.x, grp = tid >> 8, gt = tid & 255;
  if (tid < kNbr) s_off[tid] = s.g.off[tid];
  const int n2 = s.ctr[C_DSL2], n1 = s.ctr[C_DSL1];
  const int nb2 = (n2 + 255) >> 8, nb = nb2 + ((n1 + 255) >> 8);
  if (blockIdx.x == 0 && tid == 0) s.ctr[C_PCHUNK] = nb;
  const int c0 = (int)((long long)blockIdx.x * nb / gridDim.x),
            c1 = (int)((long long)(blockIdx.x + 1) * nb / gridDim.x);
  // every thread takes the same number of steps (the barriers)
  for (int cb = c0; cb < c1; cb += kG) {
    const int c = cb + grp;
    const bool live = c < c1;
    if (gt == 0) s_np[grp] = 0;
    __syncthreads();
    int2* cp = pairs + (long long)(live ? c : c0) * kChunkPairs;
    if (c < nb2) {  // uniform per group (a group's wavefronts share c)
      const int k = c * 256 + gt;
      const bool act = live && k < n2;
      saddle_entry<2>(s, kModeDec, f, s.dDS, s.dlow, s.df2, s.dl2, act ? s.dsl2[k] : 0, act,
                      s_off, cp, &s_np[grp], kmax, 1, sa, sc2, sw2);
    } else {
      const int k = (c - nb2) * 256 + gt;
      const bool act = live && k < n1;
      saddle_entry<1>(s, kModeDec, f, s.dAS, s.dlow, s.df1, s.dl1, act ? s.dsl1[k] : 0, act,
                      s_off, cp, &s_np[grp], kmin, 1, sa, sc1, sw1);
    }
    __syncthreads();
    if (gt == 0 && live) pcnt[c] = s_np[grp];
    // (a side pushes at most one entry to each list)
    sa.drain(NT);
    sc2.drain(NT);
    sc1.drain(NT);
    sw2.drain(NT);
    sw1.drain(NT);
  }
  sa.flush();
  sc2.flush();
  sc1.flush();
  sw2.flush();
  sw1.flush();
}
__global__ void __launch_bounds__(256, 4) k_inc_saddles(St s, const double* f, int2* pairs,
                                                        int* pcnt, long long* kmax,
                                                        long long* kmin) {
  inc_saddles_body<256>(s, f, pairs, pcnt, kmax, kmin);
}

// The queued slots' gathers: kPullGroup lanes per slot, a wavefront for a
// slot of more than kPullShort adjacencies (taken after its block's short
// ones; the loop's trip count is the same for every thread of a block of NT
// threads).
template <int NT>
__device__ __forceinline__ void inc_pull_body(const St& s, long long* pkey, int* pid) {
  constexpr int kPer = NT / kPullGroup;  // slots per block and step
  __shared__ int lng[kPer];
  __shared__ int nl;
  if (threadIdx.x == 0) nl = 0;
  __syncthreads();
  const int n = s.ctr[C_DSLOT];
  const int gl = threadIdx.x & (kPullGroup - 1), lane = threadIdx.x & 63;
  for (int w0 = blockIdx.x * kPer; w0 < n; w0 += gridDim.x * kPer) {
    const int w = w0 + threadIdx.x / kPullGroup;
    int j = 0, k0 = 0, k1 = 0;
    if (w < n) {
      j = s.dslot[w];
      k0 = s.csr_off[j];
      k1 = s.csr_off[j + 1];
    }
    const bool mine = w < n && k1 - k0 <= kPullShort;
    if (w < n && !mine && gl == 0) lng[atomicAdd(&nl, 1)] = j;
    bool mx = j < s.nmax;
    long long bk = 0;
    int bp = -1;
    if (mine) pull_range(s, mx, k0 + gl, k1, kPullGroup, bk, bp);
    pull_group(s, mx, kPullGroup, bk, bp);
    if (mine && gl == 0) pull_store(s, j, mx, bk, bp, pkey, pid);
    __syncthreads();
    const int nlong = nl;
    for (int q = threadIdx.x >> 6; q < nlong; q += NT / 64) {  // a wavefront per long slot
      const int jj = lng[q];
      mx = jj < s.nmax;
      bk = 0;
      bp = -1;
      pull_range(s, mx, s.csr_off[jj] + lane, s.csr_off[jj + 1], 64, bk, bp);
      pull_group(s, mx, 64, bk, bp);
      if (lane == 0) pull_store(s, jj, mx, bk, bp, pkey, pid);
    }
    __syncthreads();
    if (threadIdx.x == 0) nl = 0;
    __syncthreads();
  }
}
__global__ void __launch_bounds__(256) k_inc_pull(St s, long long* pkey, int* pid) {
  inc_pull_body<256>(s, pkey, pid);
}

// End of every CP pass of decp: clears the pass's per-vertex flags through
// its lists.
__device__ __forceinline__ void inc_clear_body(const St& s) {
  const int nu = s.ctr[C_CU], nd = s.ctr[C_CD], nl = s.ctr[C_LCHG];
  GRID_STRIDE(k, nu) {
    s.cflag[s.cu[k]] = 0;
    s.tu[s.ou[k]] = 0;
  }
  GRID_STRIDE(k, nd) {
    s.cflag[s.cd[k]] = 0;
    s.td[s.od[k]] = 0;
  }
  GRID_STRIDE(k, nl) s.lflag[s.lchg[k]] = 0;
}
__global__ void k_inc_clear(St s) {
  inc_clear_body(s);
}

// End of a CP pass of decp: the pass counters start at zero, the changed
// sides and pushed slots become the last pass's, decp counts as unchanged.
__device__ __forceinline__ void inc_end_ctr_body(int* ctr) {
  for (int i = C_LCHG + (int)threadIdx.x; i <= C_SWEEP; i += blockDim.x) ctr[i] = 0;
  if (threadIdx.x == 0) {
    ctr[C_CL2] = ctr[C_CLN2];
    ctr[C_CL1] = ctr[C_CLN1];
    ctr[C_PPREV] = ctr[C_PNEXT];
    ctr[C_PNEXT] = 0;
    ctr[C_ANY_CHANGE] = 0;
  }
}
__global__ void k_inc_end_ctr(int* ctr) {
  inc_end_ctr_body(ctr);
}

// A phase boundary inside one block whose phases hand lists and counters
// to each other through global memory (all of it through this CU's L2): every
// wavefront's loads, stores and atomics have completed before the barrier,
// and the acquire after it drops the lines the CU's vector cache may hold
// from before them (the counters are appended to with atomics, which that
// cache does not see).
__device__ __forceinline__ void block_phase() {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// The whole incremental CP pass in one block of 1024 threads, for passes of
// few changed vertices (every list of such a pass holds a few thousand
// entries at most): the phases of k_inc_ptrs ... k_slot_ids in order, the
// root propagation entirely level by level in the block (no whole-grid
// sweeps), and with end the end-of-pass clears of k_inc_clear and
// k_inc_end_ctr.  Not for the first pass after a full one (every slot is
// merged then).  Same results as the multi-block sequence: each phase is the
// same code over the same lists, in the same order.
__global__ void __launch_bounds__(1024) k_inc_one(St s, int epoch, long long* kmax,
                                                  long long* kmin, int2* pairs, int* pcnt,
                                                  long long* pkey, int* pid, int end) {
  inc_ptrs_body(s, epoch);
  block_phase();
  inc_taint_body(s);
  block_phase();
  inc_walk_body(s);
  block_phase();
  inc_prop_body(s, s.g.N);
  block_phase();
  inc_mark_body(s, kmax, kmin, 0);
  block_phase();
  inc_mark_csr_body(s);
  block_phase();
  inc_saddles_body<1024>(s, s.decp, pairs, pcnt, kmax, kmin);
  block_phase();
  inc_pull_body<1024>(s, pkey, pid);
  block_phase();
  inc_merge_body(s, kmax, kmin, pkey, pid, pairs, pcnt, 0);
  block_phase();
  slot_ids_body(s, pairs, pcnt, s.decp, kmax, kmin, s.dlarg, s.dsmall);
  if (!end) return;
  block_phase();
  inc_clear_body(s);
  block_phase();
  inc_end_ctr_body(s.ctr);
}

// Number of differing bytes of two buffers (out[0], accumulated).
__global__ void k_diff_count(const void* a, const void* b, long long nbytes, int* out) {
  const uint8_t* x = (const uint8_t*)a;
  const uint8_t* y = (const uint8_t*)b;
  int c = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nbytes;
       i += (long long)gridDim.x * blockDim.x)
    c += x[i] != y[i];
  if (c) atomicAdd(out, c);
}

// ------------------------------------------- wrong neighbours (greedy) ----
// Deterministic reservations reproducing the serial greedy of
// get_wrong_split_neighbors / get_wrong_join_neighbors in saddle-id order
// (main.cpp:309-323).  Entries are {2*saddle + side, candidate mask}; side 0
// is a 2-saddle (upper star), side 1 a 1-saddle (lower star).
__global__ void k_act_copy(St s, int2* dst, int dctr) {
  int n = s.ctr[C_ACT0];
  GRID_STRIDE(i, n) dst[i] = s.act0[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) s.ctr[dctr] = n;
}

// Drop flagged candidates; every remaining candidate records the smallest
// saddle id that may still take it.
__global__ void __launch_bounds__(256) k_wn_register(St s, int2* active, int actr, int nctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) s.ctr[nctr] = 0;  // the list k_wn_commit fills
  int n = s.ctr[actr];
  GRID_STRIDE(a, n) {
    int2 e = active[a];
    int sv = e.x >> 1, side = e.x & 1;
    const uint8_t* fl = side ? s.wn1 : s.wn2;
    int* ms = side ? s.minsad1 : s.minsad2;
    unsigned mask = (unsigned)e.y;
    for (int d = 0; d < kNbr; d++) {
      if (!((mask >> d) & 1)) continue;
      int r = sv + s.g.off[d];
      if (fl[r]) {
        mask &= ~(1u << d);
        continue;
      }
      atomicMin(&ms[r], sv);
    }
    active[a].y = (int)mask;
  }
}

// A saddle whose first candidate is reserved by itself takes it (flag +
// append, as get_wrong_split_neighbors does before its return); the others
// go to the next round's list.  Each saddle also releases the reservations it
// holds (those equal to its id): a reservation is changed only by its holder,
// and any other value a saddle reads there (the holder's id, or the released
// one) tells it the same, that it does not hold it.
__global__ void __launch_bounds__(256) k_wn_commit(St s, const int2* active, int actr,
                                                   int2* next, int nctr) {
  STAGED_LIST(l2, int, kStage, s.wnl2, &s.ctr[C_NFC]);
  STAGED_LIST(l1, int, kStage, s.wnl1, &s.ctr[C_NFC1]);
  STAGED_LIST(nl, int2, kStage, next, &s.ctr[nctr]);
  int n = s.ctr[actr];
  GRID_STRIDE(a, n) {
    const int2 e = active[a];
    const unsigned mask = (unsigned)e.y;
    if (!mask) continue;
    const int sv = e.x >> 1, side = e.x & 1;
    int* ms = side ? s.minsad1 : s.minsad2;
    const int first = __ffs(mask) - 1;
    bool took = false;
    for (int d = first; d < kNbr; d++) {
      if (!((mask >> d) & 1)) continue;
      const int r = sv + s.g.off[d];
      if (ms[r] == sv) {
        ms[r] = 0x7fffffff;
        took |= d == first;
      }
    }
    if (took) {
      const int r = sv + s.g.off[first];
      if (!side) {
        s.wn2[r] = 1;
        l2.push(r);
      } else {
        s.wn1[r] = 1;
        l1.push(r);
      }
    } else {
      nl.push(e);
    }
  }
  l2.flush();
  l1.flush();
  nl.flush();
}

// The rounds of k_wn_solve with every reservation, flag and active entry in
// LDS, for candidate lists of at most kWnLds entries whose distinct
// candidate vertices fill at most half of a kWnHash-slot table (open
// addressing on (vertex, side)).  Returns false, having written nothing to
// global memory, when the list does not fit; the caller then runs the global
// rounds.  The committed wrong neighbours go to wnl2 / wnl1 (the vertex flags
// wn2 / wn1 are not needed: k_wn_solve and k_wn_reset clear them through
// those lists).  Called by the whole block of 1024 threads.
constexpr int kWnLds = 4096, kWnHash = 8192;
constexpr unsigned kWnEmpty = 0xffffffffu;

__device__ __forceinline__ unsigned wn_slot0(unsigned key) {
  return (key * 2654435761u) >> (32 - 13);  // kWnHash = 2^13
}
// The slot of a key known to be in the table.
__device__ __forceinline__ int wn_find(const unsigned* keys, unsigned key) {
  unsigned h = wn_slot0(key);
  while (keys[h] != key) h = (h + 1) & (kWnHash - 1);
  return (int)h;
}

__device__ bool wn_solve_lds(const St& s) {
  __shared__ unsigned hkey[kWnHash];
  __shared__ int hmin[kWnHash];
  __shared__ uint8_t hflag[kWnHash];
  __shared__ int2 la[kWnLds], lb[kWnLds];
  __shared__ int n_cur, n_next, n_fc, n_fc1, nkeys, ok;
  const int t = threadIdx.x;
  if (t == 0) {
    n_cur = s.ctr[C_ACT0];
    n_next = 0;
    n_fc = 0;
    n_fc1 = 0;
    nkeys = 0;
    ok = n_cur <= kWnLds;
  }
  __syncthreads();
  if (!ok) return false;
  const int n0 = n_cur;
  for (int i = t; i < kWnHash; i += 1024) {
    hkey[i] = kWnEmpty;
    hmin[i] = 0x7fffffff;
    hflag[i] = 0;
  }
  for (int i = t; i < n0; i += 1024) la[i] = s.act0[i];
  __syncthreads();
  // every candidate (vertex, side) once in the table
  for (int a = t; a < n0; a += 1024) {
    const int2 e = la[a];
    const int sv = e.x >> 1, side = e.x & 1;
    unsigned mask = (unsigned)e.y;
    while (mask) {
      const int d = __ffs(mask) - 1;
      mask &= mask - 1;
      const unsigned key = ((unsigned)(sv + s.g.off[d]) << 1) | (unsigned)side;
      unsigned h = wn_slot0(key);
      for (int p = 0;; p++) {
        const unsigned old = atomicCAS(&hkey[h], kWnEmpty, key);
        if (old == kWnEmpty) {
          if (atomicAdd(&nkeys, 1) >= kWnHash / 2) ok = 0;
          break;
        }
        if (old == key) break;
        if (p == kWnHash / 2) {  // (cannot happen while the table is at most half full)
          ok = 0;
          break;
        }
        h = (h + 1) & (kWnHash - 1);
      }
    }
  }
  __syncthreads();
  if (!ok) return false;
  int2* act = la;
  int2* nxt = lb;
  for (int round = 0; n_cur > 0 && round <= n0; round++) {
    const int n = n_cur;
    for (int a = t; a < n; a += 1024) {  // register
      const int2 e = act[a];
      const int sv = e.x >> 1, side = e.x & 1;
      unsigned mask = (unsigned)e.y, m = mask;
      while (m) {
        const int d = __ffs(m) - 1;
        m &= m - 1;
        const int h = wn_find(hkey, ((unsigned)(sv + s.g.off[d]) << 1) | (unsigned)side);
        if (hflag[h]) mask &= ~(1u << d);
        else atomicMin(&hmin[h], sv);
      }
      act[a].y = (int)mask;
    }
    __syncthreads();
    for (int a = t; a < n; a += 1024) {  // commit
      const int2 e = act[a];
      const unsigned mask = (unsigned)e.y;
      if (!mask) continue;
      const int sv = e.x >> 1, side = e.x & 1;
      const int r = sv + s.g.off[__ffs(mask) - 1];
      const int h = wn_find(hkey, ((unsigned)r << 1) | (unsigned)side);
      if (hmin[h] == sv) {
        hflag[h] = 1;
        if (!side) s.wnl2[atomicAdd(&n_fc, 1)] = r;
        else s.wnl1[atomicAdd(&n_fc1, 1)] = r;
        act[a].y = (int)(mask | 0x8000u);  // done
      }
    }
    __syncthreads();
    for (int a = t; a < n; a += 1024) {  // release and compaction
      const int2 e = act[a];
      const unsigned mask = (unsigned)e.y;
      const int sv = e.x >> 1, side = e.x & 1;
      unsigned m = mask & ((1u << kNbr) - 1);
      while (m) {
        const int d = __ffs(m) - 1;
        m &= m - 1;
        hmin[wn_find(hkey, ((unsigned)(sv + s.g.off[d]) << 1) | (unsigned)side)] = 0x7fffffff;
      }
      if (mask && !(mask & 0x8000u)) nxt[atomicAdd(&n_next, 1)] = e;
    }
    __syncthreads();
    if (t == 0) {
      n_cur = n_next;
      n_next = 0;
    }
    int2* tmp = act;
    act = nxt;
    nxt = tmp;
    __syncthreads();
  }
  if (t == 0) {
    s.ctr[C_NFC] = n_fc;
    s.ctr[C_NFC1] = n_fc1;
    if (n_cur > 0) s.ctr[C_WN_BAIL] = 2;  // every round commits one saddle: unreachable
  }
  return true;
}

// All rounds of the reservations above in one block of 1024 threads, for
// candidate lists of at most `limit` entries (else C_WN_BAIL is set and
// nothing is done): the flag reset of k_wn_reset, the copy of k_act_copy and
// the register / commit / next phases, separated by block barriers (which
// order the block's plain accesses).  The reservations are atomics, which the
// CU's L1 does not see: they are read and cleared with atomic loads and
// stores.
__global__ void __launch_bounds__(1024) k_wn_solve(St s, int2* bufA, int2* bufB, int limit) {
  __shared__ int n_cur, n_next, n_fc, n_fc1, bail;
  const int t = threadIdx.x;
  if (t == 0) {
    const int n0 = s.ctr[C_ACT0];
    bail = n0 > limit;
    s.ctr[C_WN_BAIL] = bail;
    n_cur = n0;
    n_next = 0;
    n_fc = 0;
    n_fc1 = 0;
  }
  __syncthreads();
  if (bail) return;
  uint8_t* wn2 = s.wn2;
  uint8_t* wn1 = s.wn1;
  {
    const int n = s.ctr[C_NFC], n1 = s.ctr[C_NFC1];
    for (int i = t; i < n; i += 1024) wn2[s.wnl2[i]] = 0;
    for (int i = t; i < n1; i += 1024) wn1[s.wnl1[i]] = 0;
  }
  if (wn_solve_lds(s)) return;
  for (int i = t; i < n_cur; i += 1024) bufA[i] = s.act0[i];
  __syncthreads();
  int2* act = bufA;
  int2* nxt = bufB;
  for (int round = 0; n_cur > 0 && round <= limit; round++) {
    const int n = n_cur;
    int2* va = act;
    // register (k_wn_register)
    for (int a = t; a < n; a += 1024) {
      const int ex = va[a].x;
      unsigned mask = (unsigned)va[a].y;
      const int sv = ex >> 1, side = ex & 1;
      const uint8_t* fl = side ? wn1 : wn2;
      int* ms = side ? s.minsad1 : s.minsad2;
      for (int d = 0; d < kNbr; d++) {
        if (!((mask >> d) & 1)) continue;
        const int r = sv + s.g.off[d];
        if (fl[r]) {
          mask &= ~(1u << d);
          continue;
        }
        atomicMin(&ms[r], sv);
      }
      va[a].y = (int)mask;
    }
    __syncthreads();
    // commit (k_wn_commit)
    for (int a = t; a < n; a += 1024) {
      const int ex = va[a].x;
      const unsigned mask = (unsigned)va[a].y;
      if (!mask) continue;
      const int sv = ex >> 1, side = ex & 1;
      const int r = sv + s.g.off[__ffs(mask) - 1];
      int* ms = side ? s.minsad1 : s.minsad2;
      if (__hip_atomic_load(&ms[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == sv) {
        if (!side) {
          wn2[r] = 1;
          s.wnl2[atomicAdd(&n_fc, 1)] = r;
        } else {
          wn1[r] = 1;
          s.wnl1[atomicAdd(&n_fc1, 1)] = r;
        }
        va[a].y = (int)(mask | 0x8000u);  // done
      }
    }
    __syncthreads();
    // release and compaction
    for (int a = t; a < n; a += 1024) {
      const int ex = va[a].x;
      const unsigned mask = (unsigned)va[a].y;
      const int sv = ex >> 1, side = ex & 1;
      int* ms = side ? s.minsad1 : s.minsad2;
      for (int d = 0; d < kNbr; d++)
        if ((mask >> d) & 1)
          __hip_atomic_store(&ms[sv + s.g.off[d]], 0x7fffffff, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
      if (mask && !(mask & 0x8000u)) nxt[atomicAdd(&n_next, 1)] = make_int2(ex, (int)mask);
    }
    __syncthreads();
    if (t == 0) {
      n_cur = n_next;
      n_next = 0;
    }
    int2* tmp = act;
    act = nxt;
    nxt = tmp;
    __syncthreads();
  }
  if (t == 0) {
    s.ctr[C_NFC] = n_fc;
    s.ctr[C_NFC1] = n_fc1;
    if (n_cur > 0) s.ctr[C_WN_BAIL] = 2;  // every round commits one saddle: unreachable
  }
}

__global__ void k_wn_reset(St s) {
  int n = s.ctr[C_NFC];
  GRID_STRIDE(i, n) s.wn2[s.wnl2[i]] = 0;
  int n1 = s.ctr[C_NFC1];
  GRID_STRIDE(i, n1) s.wn1[s.wnl1[i]] = 0;
}

// ------------------------------------------------------------ fixpath ----
// fixpath, split (src/ExtreM.cpp:2452-2575): walk the original steepest path
// from a wrong neighbour until decp's steepest choice differs.
// With fast, decp has not changed since its last CP pass, so the steepest
// directions of both fields are the stored ones (opdir, pdir) and each step
// is two byte loads instead of the 28 values of the stencil.
__global__ void __launch_bounds__(256) k_fixpath(St s, int fast) {
  STAGED_LIST(tl, int, kStage, s.touched, &s.ctr[s.tctr]);
  int n2 = s.ctr[C_NFC], n1 = s.ctr[C_NFC1];
  block_rounds(n2 + n1, 1, [&](long long ee) {
    const int e = (int)ee;
    int dir = e < n2 ? 0 : 1;
    int cur = dir == 0 ? s.wnl2[e] : s.wnl1[e - n2];
    int tru = kNone, fal = kNone;
    if (fast) {
      const int sh = dir == 0 ? 0 : 4;
      for (int step = 0; step <= s.g.N; step++) {
        const int a = (s.opdir[cur] >> sh) & 15, b = (s.pdir[cur] >> sh) & 15;
        const int nx = a ? cur + s.g.off[a - 1] : cur, dn = b ? cur + s.g.off[b - 1] : cur;
        if (nx != dn) {
          tru = nx;
          fal = dn;
          break;
        }
        if (nx == cur) break;
        cur = nx;
      }
    } else
    for (int step = 0; step <= s.g.N; step++) {
      int x, y, z;
      coords(s.g, cur, x, y, z);
      int key = c_cls_key[class_of(s.g, x, y, z)];
      int nx = cur, dn = cur;
      double fo = s.orig[cur], fd = s.decp[cur];
      for (int d = 0; d < kNbr; d++) {
        if (!((key >> d) & 1)) continue;
        int r = cur + s.g.off[d];
        double ro = s.orig[r], rd = s.decp[r];
        if (dir == 0) {
          if (fo < ro || (fo == ro && nx < r)) { nx = r; fo = ro; }
          if (fd < rd || (fd == rd && dn < r)) { dn = r; fd = rd; }
        } else {
          if (fo > ro || (fo == ro && nx > r)) { nx = r; fo = ro; }
          if (fd > rd || (fd == rd && dn > r)) { dn = r; fd = rd; }
        }
      }
      if (nx != dn) {
        tru = nx;
        fal = dn;
        break;
      }
      if (nx == cur) break;
      cur = nx;
    }
    if (tru == fal) return;
    if (dir == 0) propose(s, fal, lower_mid(s, fal), tl);
    else propose(s, tru, lower_mid(s, tru), tl);
  }, tl);
  tl.flush();
}

// ------------------------------------------------- wrong index (ranks) ----
// get_wrong_index_max / get_wrong_index_min (src/ExtreM.cpp:2595-2670): the
// first saddle (lowest id) per original extremum wins; pairs are compared by
// vertex id.  phase 0 registers, phase 1 counts (and proposes, s_loops /
// s_loops_join, src/ExtreM.cpp:2838-2886) and releases: a registration word
// holds its winner's id until the winner itself resets it, and any other
// value a loser reads (that id or the reset one) tells it the same.
// The registration words (win_a / win_b) are one per listed extremum, at its
// slot: an original triplet end is an extremum of the original field, so it
// has a list position.
__device__ __forceinline__ int end_slot(const St& s, int e) { return slot_index(s, s.ext_id[e]); }

__global__ void __launch_bounds__(256) k_wrong_index(St s, int side, int phase, int do_fix) {
  // only the listed sides can have triplet ends that differ
  const int n = s.ctr[side == 2 ? C_WL2 : C_WL1];
  const int* WL = side == 2 ? s.wl2 : s.wl1;
  const int* S = side == 2 ? s.S2 : s.S1;
  const int* of = side == 2 ? s.of2 : s.of1;
  const int* ol = side == 2 ? s.ol2 : s.ol1;
  const int* df = side == 2 ? s.df2 : s.df1;
  const int* dl = side == 2 ? s.dl2 : s.dl1;
  STAGED_LIST(tl, int, kStage, s.touched, &s.ctr[s.tctr]);
  STAGED_COUNT(cf, &s.ctr[side == 2 ? C_WMC : C_WMINC2]);
  STAGED_COUNT(cl, &s.ctr[side == 2 ? C_WMC2 : C_WMINC]);
  if (phase == 0 && blockIdx.x == 0 && threadIdx.x == 0) {  // counted by phase 1
    s.ctr[side == 2 ? C_WMC : C_WMINC2] = 0;
    s.ctr[side == 2 ? C_WMC2 : C_WMINC] = 0;
  }
  // (at most two proposals per listed side)
  block_rounds(n, 2, [&](long long kk) {
    const int k = (int)kk;
    const int i = WL[k];
    int sv = S[i];
    int a = of[i], b = ol[i], c = df[i], d = dl[i];
    // a NONE end (an empty triplet: only the persistence filter makes one)
    // is no mismatch on either side
    bool mf = a != c && a != kNone && c != kNone, ml = b != d && b != kNone && d != kNone;
    if (phase == 0) {
      if (mf) atomicMin(&s.win_a[end_slot(s, a)], sv);
      if (ml) atomicMin(&s.win_b[end_slot(s, b)], sv);
    } else {
      if (mf && s.win_a[end_slot(s, a)] == sv) {
        s.win_a[end_slot(s, a)] = 0x7fffffff;
        cf.inc();
        if (side == 2) {  // triplet[0] -> wrong_max_counter, fix dir 0 lowers orig
          if (do_fix) propose(s, a, lower_mid(s, a), tl);
        } else {          // triplet[0] -> wrong_min_counter_2, fix dir 1 lowers decp's
          if (do_fix && c != kNone) propose(s, c, lower_mid(s, c), tl);
        }
      }
      if (ml && s.win_b[end_slot(s, b)] == sv) {
        s.win_b[end_slot(s, b)] = 0x7fffffff;
        cl.inc();
        if (side == 2) {  // triplet[last] -> wrong_max_counter_2, fix dir 1 lowers decp's
          if (do_fix && d != kNone) propose(s, d, lower_mid(s, d), tl);
        } else {          // triplet[last] -> wrong_min_counter, fix dir 0 lowers orig
          if (do_fix) propose(s, b, lower_mid(s, b), tl);
        }
      }
    }
  }, tl);
  tl.flush();
  cf.flush();
  cl.flush();
}

// The three phases of k_wrong_index in one block of 1024 threads, for listed
// sides the host knows to be few: block barriers order the phases, and the
// registrations are atomics, which the CU's L1 does not see, so they are read
// and released with atomic loads and stores.  With wsc >= 0 the block, whose
// counts are then final, also opens saddle_loops' gate as k_saddle_gate does.
__global__ void __launch_bounds__(1024) k_wrong_index_one(St s, int side, int do_fix, int wsc) {
  const int n = s.ctr[side == 2 ? C_WL2 : C_WL1];
  const int* WL = side == 2 ? s.wl2 : s.wl1;
  const int* S = side == 2 ? s.S2 : s.S1;
  const int* of = side == 2 ? s.of2 : s.of1;
  const int* ol = side == 2 ? s.ol2 : s.ol1;
  const int* df = side == 2 ? s.df2 : s.df1;
  const int* dl = side == 2 ? s.dl2 : s.dl1;
  STAGED_LIST(tl, int, kStage, s.touched, &s.ctr[s.tctr]);
  STAGED_COUNT(cf, &s.ctr[side == 2 ? C_WMC : C_WMINC2]);
  STAGED_COUNT(cl, &s.ctr[side == 2 ? C_WMC2 : C_WMINC]);
  const int t = threadIdx.x;
  if (t == 0) {  // counted below (the flushes come after, from this thread)
    s.ctr[side == 2 ? C_WMC : C_WMINC2] = 0;
    s.ctr[side == 2 ? C_WMC2 : C_WMINC] = 0;
  }
  // a NONE end (an empty triplet) is no mismatch, as in k_wrong_index
  auto differs = [](int o, int dv) { return o != dv && o != kNone && dv != kNone; };
  for (int k = t; k < n; k += 1024) {  // register
    const int i = WL[k];
    const int sv = S[i], a = of[i], b = ol[i];
    if (differs(a, df[i])) atomicMin(&s.win_a[end_slot(s, a)], sv);
    if (differs(b, dl[i])) atomicMin(&s.win_b[end_slot(s, b)], sv);
  }
  __syncthreads();
  for (int k = t; k < n; k += 1024) {  // count and fix
    const int i = WL[k];
    const int sv = S[i];
    const int a = of[i], b = ol[i], c = df[i], d = dl[i];
    if (differs(a, c) &&
        __hip_atomic_load(&s.win_a[end_slot(s, a)], __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT) == sv) {
      cf.inc();
      if (side == 2) {
        if (do_fix) propose(s, a, lower_mid(s, a), tl);
      } else {
        if (do_fix && c != kNone) propose(s, c, lower_mid(s, c), tl);
      }
    }
    if (differs(b, d) &&
        __hip_atomic_load(&s.win_b[end_slot(s, b)], __ATOMIC_RELAXED,
                          __HIP_MEMORY_SCOPE_AGENT) == sv) {
      cl.inc();
      if (side == 2) {
        if (do_fix && d != kNone) propose(s, d, lower_mid(s, d), tl);
      } else {
        if (do_fix) propose(s, b, lower_mid(s, b), tl);
      }
    }
  }
  __syncthreads();
  for (int k = t; k < n; k += 1024) {  // release
    const int i = WL[k];
    const int a = of[i], b = ol[i];
    if (differs(a, df[i]))
      __hip_atomic_store(&s.win_a[end_slot(s, a)], 0x7fffffff, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    if (differs(b, dl[i]))
      __hip_atomic_store(&s.win_b[end_slot(s, b)], 0x7fffffff, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  tl.flush();
  cf.flush();
  cl.flush();
  if (wsc >= 0 && t == 0) {  // the flushes' adds were this thread's own
    const int gate = side == 2 ? C_WMC : C_WMINC;
    if (__hip_atomic_load(&s.ctr[gate], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
      s.ctr[wsc] = 0;
      s.ctr[C_SLC]++;
    }
  }
}

// get_wrong_index_saddles(_join) + fix_wrong_index_saddle
// (src/ExtreM.cpp:2672-2715, 2811-2836).  The global extremum has no saddle
// (its triplet entries are skipped, src/ExtreM.cpp:2330); a side without an
// adjacent saddle (the reference reads one past its saddle vector there) is
// no mismatch.
// saddle_loops' condition (main.cpp:250-251, 270-271, 328-336, 378-387): the side's
// wrong-index count is zero.  Then its wrong-saddle count starts at zero and
// the call is counted; else k_wrong_saddle with this gate does nothing.
__global__ void k_saddle_gate(int* ctr, int gate, int wsc) {
  if (threadIdx.x == 0 && blockIdx.x == 0 && ctr[gate] == 0) {
    ctr[wsc] = 0;
    ctr[C_SLC]++;
  }
}

__global__ void __launch_bounds__(256) k_wrong_saddle(St s, int side, int do_fix, int gate) {
  if (gate >= 0 && s.ctr[gate] != 0) return;  // the whole grid reads the same count
  int n = side == 2 ? s.nmax : s.nmin;
  const int* E = side == 2 ? s.MX : s.MN;
  const int* os = side == 2 ? s.olarg : s.osmall;
  const int* ds = side == 2 ? s.dlarg : s.dsmall;
  int og = s.ctr[side == 2 ? C_GMAX : C_GMIN];
  int dg = s.ctr[side == 2 ? C_DGMAX : C_DGMIN];
  STAGED_LIST(tl, int, kStage, s.touched, &s.ctr[s.tctr]);
  STAGED_COUNT(cnt, &s.ctr[side == 2 ? C_WSC : C_WSCJ]);
  block_rounds(n, 1, [&](long long k) {
    const int i = (int)k;
    int m = E[i];
    int ov = m == og ? kNone : os[i];
    int dv = m == dg ? kNone : ds[i];
    if (ov == 0x7fffffff) ov = kNone;  // empty minimum slot
    if (dv == 0x7fffffff) dv = kNone;
    if (ov == dv || ov == kNone || dv == kNone) return;
    cnt.inc();
    if (!do_fix) return;
    double ft = s.decp[ov], ff = s.decp[dv];
    if (ft > ff || (ft == ff && ov > dv)) return;
    if (side == 2) propose(s, dv, lower_mid(s, dv), tl);
    else propose(s, ov, raise_mid(s, ov), tl);
  }, tl);
  tl.flush();
  cnt.flush();
}

// --------------------------------------------------------------- setup ----
// The saddle / extremum lists by original type, in an ordered compaction
// (per-block counts, one scan, ordered writes) over a traversal of the grid
// in columns of kListTile x kListTile voxels, each column marched along z
// (for fields at least that wide; plain vertex order otherwise).  List
// neighbours are then grid neighbours within a few voxels in every axis, so
// the per-saddle passes reuse the neighbour rows of a column in L2 instead
// of streaming whole planes.  No pass depends on the list order: every
// order-dependent choice compares vertex ids.
constexpr int kListTile = 32;

// The traversal in units of 1024 positions, kListTile rows of kListTile: one
// z-plane of one column (unit col * D + z, col = cy * ntx + cx) in the tiled
// order, 1024 consecutive vertices otherwise.  A half-wavefront takes one
// unit, a block of 256 threads eight: in the tiled order eight neighbouring
// columns of one z-plane (block (cy * ngx + gx) * D + z), so every row the
// block reads is 256 consecutive voxels of the grid.
struct UnitLane {
  int u;  // the unit (-1: a column past the grid's last)
  int x0, y0, z;
};
template <bool TILED>
__device__ __forceinline__ UnitLane unit_lane(const Geo& g, int ntx, int ngx) {
  const int hw = threadIdx.x >> 5;
  UnitLane ul{0, 0, 0, 0};
  if (TILED) {
    const int b = blockIdx.x;
    ul.z = b % g.D;
    const int q = b / g.D;
    const int cx = (q % ngx) * 8 + hw, cy = q / ngx;
    ul.x0 = cx * kListTile;
    ul.y0 = cy * kListTile;
    ul.u = cx < ntx ? (cy * ntx + cx) * g.D + ul.z : -1;
  } else {
    ul.u = blockIdx.x * 8 + hw;
  }
  return ul;
}
// The vertex at row r of the unit, or -1 for a position outside the grid.
template <bool TILED>
__device__ __forceinline__ int unit_vertex(const Geo& g, const UnitLane& ul, int r) {
  const int l = threadIdx.x & 31;
  if (TILED) {
    const int x = ul.x0 + l, y = ul.y0 + r;
    return (ul.u >= 0 && x < g.W && y < g.H) ? x + g.W * (y + g.H * ul.z) : -1;
  }
  const long long k = (long long)ul.u * (kListTile * kListTile) + r * kListTile + l;
  return k < g.N ? (int)k : -1;
}
// The original type at v, loaded for every lane (at vertex 0 for v < 0), so
// the loads of successive rows need no branch and stay in flight together.
__device__ __forceinline__ int unit_type(const St& s, int v) { return s.otype[v >= 0 ? v : 0]; }
// List flags of a vertex by its original type t (0 for v < 0).
__device__ __forceinline__ unsigned unit_flags(int v, int t) {
  const unsigned f = (unsigned)(t == 2 || t == 3) | (unsigned)(t == 1 || t == 3) << 1 |
                     (unsigned)(t == 4) << 2 | (unsigned)(t == 0) << 3;
  return v >= 0 ? f : 0u;
}

// cnt[c * nb + unit] = number of category-c vertices in the unit
// (categories 2-saddle, 1-saddle, maximum, minimum; nb units).
template <bool TILED>
__global__ void __launch_bounds__(256) k_lists_count(St s, int* cnt, int nb, int ntx, int ngx) {
  const UnitLane ul = unit_lane<TILED>(s.g, ntx, ngx);
  int n[4] = {0, 0, 0, 0};
#pragma unroll 8
  for (int r = 0; r < kListTile; r++) {
    const int v = unit_vertex<TILED>(s.g, ul, r);
    const unsigned fl = unit_flags(v, unit_type(s, v));
#pragma unroll
    for (int c = 0; c < 4; c++) n[c] += (fl >> c) & 1u;
  }
#pragma unroll
  for (int c = 0; c < 4; c++)
    for (int o = 16; o > 0; o >>= 1) n[c] += __shfl_xor(n[c], o);  // within the half
  if ((threadIdx.x & 31) == 0 && ul.u >= 0 && ul.u < nb) {
#pragma unroll
    for (int c = 0; c < 4; c++) cnt[(long long)c * nb + ul.u] = n[c];
  }
}
template __global__ void k_lists_count<true>(St, int*, int, int, int);
template __global__ void k_lists_count<false>(St, int*, int, int, int);

// List lengths from the exclusive scan of the 4 * nb block counts taken as
// one array (entry 4 * nb holds the grand total): category c's blocks start
// at cnt[c * nb], so its length is cnt[(c + 1) * nb] - cnt[c * nb].
__global__ void k_lists_totals(St s, const int* cnt, int nb) {
  if (blockIdx.x == 0 && threadIdx.x < 4) {
    const int c = threadIdx.x;
    s.ctr[C_N2 + c] = cnt[(long long)(c + 1) * nb] - cnt[(long long)c * nb];
  }
}

// With index, every vertex also gets its list positions: spos2 / spos1 (-1
// off the saddle lists) and ext_id (2 * position in MX, 2 * position in MN +
// 1, -1 elsewhere); the traversal visits each vertex once, so nothing needs
// clearing first.
template <bool TILED>
__global__ void __launch_bounds__(256) k_lists_write(St s, const int* cnt, int nb, int ntx, int ngx,
                                                     int index) {
  const UnitLane ul = unit_lane<TILED>(s.g, ntx, ngx);
  const bool act = ul.u >= 0 && ul.u < nb;
  const int shift = threadIdx.x & 32;  // this half's bits of a ballot
  const unsigned below = (1u << (threadIdx.x & 31)) - 1u;
  int* lists[4] = {s.S2, s.S1, s.MX, s.MN};
  int base[4];  // the unit's next position per list, the same in the half's lanes
#pragma unroll
  for (int c = 0; c < 4; c++)
    base[c] = act ? cnt[(long long)c * nb + ul.u] - cnt[(long long)c * nb] : 0;
  // every row's type is loaded before the first store: a wait for a load
  // also waits for the stores issued before it
  int tv[kListTile];
#pragma unroll
  for (int r = 0; r < kListTile; r++) tv[r] = unit_type(s, act ? unit_vertex<TILED>(s.g, ul, r) : -1);
#pragma unroll
  for (int r = 0; r < kListTile; r++) {
    const int v = act ? unit_vertex<TILED>(s.g, ul, r) : -1;
    const unsigned fl = unit_flags(v, tv[r]);
    int pos[4] = {-1, -1, -1, -1};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const unsigned hm = (unsigned)(__ballot((fl >> c) & 1u) >> shift);
      if ((fl >> c) & 1u) {
        pos[c] = base[c] + __popc(hm & below);
        lists[c][pos[c]] = v;
      }
      base[c] += __popc(hm);
    }
    if (index && v >= 0) {
      s.spos2[v] = pos[0];
      s.spos1[v] = pos[1];
      s.ext_id[v] = pos[2] >= 0 ? 2 * pos[2] : (pos[3] >= 0 ? 2 * pos[3] + 1 : -1);
    }
  }
}
template __global__ void k_lists_write<true>(St, const int*, int, int, int, int);
template __global__ void k_lists_write<false>(St, const int*, int, int, int, int);

// Global minimum and maximum of f by (value, id) (main.cpp:67-75), first
// stage: each block's best into bmin / bmax[blockIdx.x]; k_ext_reduce over
// those with one block finishes (no atomics on the two result words).  With
// check, a value that is not finite (or, with f32, not a float32) sets
// ctr[C_FAIL]: the same read serves the input check.
__global__ void __launch_bounds__(256) k_argext(const double* f, long long n, int* bmin,
                                                int* bmax, int* ctr, int check, int f32) {
  int cmin = kNone, cmax = kNone;
  double fmin = 0.0, fmax = 0.0;
  bool bad = false;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double fv = f[i];
    const int v = (int)i;
    if (check) bad |= !isfinite(fv) || (f32 && (double)(float)fv != fv);
    if (cmin == kNone || smaller(fv, v, fmin, cmin)) { cmin = v; fmin = fv; }
    if (cmax == kNone || larger(fv, v, fmax, cmax)) { cmax = v; fmax = fv; }
  }
  if (bad) ctr[C_FAIL] = 1;
  __shared__ int s_best[4];
  const int gm = block_best(cmax, fmax, f, true, s_best);
  const int gn = block_best(cmin, fmin, f, false, s_best);
  if (threadIdx.x == 0) {
    bmax[blockIdx.x] = gm;
    bmin[blockIdx.x] = gn;
  }
}

// Number of vertices with a non-zero edit count: 4 counts per load where the
// array is 4-byte aligned, one atomic per block.
__global__ void __launch_bounds__(256) k_count_edited(const uint8_t* cnt, long long n, int* ctr) {
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long nt = (long long)gridDim.x * blockDim.x;
  int c = 0;
  long long done = 0;
  if (((uintptr_t)cnt & 3) == 0) {
    const unsigned* w = (const unsigned*)cnt;
    const long long nw = n >> 2;
    for (long long i = tid; i < nw; i += nt) {
      const unsigned x = w[i];
      c += ((x & 0xffu) != 0) + ((x & 0xff00u) != 0) + ((x & 0xff0000u) != 0) +
           ((x & 0xff000000u) != 0);
    }
    done = nw << 2;
  }
  for (long long i = done + tid; i < n; i += nt) c += cnt[i] != 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  __shared__ int s_c[4];
  if (__lane_id() == 0) s_c[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int t = s_c[0] + s_c[1] + s_c[2] + s_c[3];
    if (t) atomicAdd(&ctr[C_EDITED], t);
  }
}

// --------------------------------------------------------- encoding -----
// 3-bit LSB-first packing (encodeTo3BitBitmap, src/ExtreM.cpp:220-250): byte
// j holds bits [8j, 8j+8) of the stream where value i occupies bits
// [3i, 3i+3).  One byte per lane; values are 0 where decp is unchanged.
__global__ void k_bitmap3(const double* before, const double* after, const uint8_t* cnt,
                          long long n, uint8_t* out, long long nbytes) {
  for (long long j = blockIdx.x * (long long)blockDim.x + threadIdx.x; j < nbytes;
       j += (long long)gridDim.x * blockDim.x) {
    long long b0 = 8 * j;
    long long i0 = b0 / 3, i1 = (b0 + 7) / 3;
    unsigned acc = 0;
    for (long long i = i0; i <= i1 && i < n; i++) {
      unsigned v = before[i] != after[i] ? cnt[i] : 0u;
      long long sh = 3 * i - b0;
      acc |= sh >= 0 ? (v << sh) : (v >> (-sh));
    }
    out[j] = (uint8_t)(acc & 0xffu);
  }
}

// Edits of snapped vertices in index order: flag pass + block scan is done
// on the host side of the call; this writes edit i to position pos[i].
__global__ void k_edit_flags(const double* before, const double* after, const uint8_t* cnt,
                             long long n, int* flag) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    flag[i] = (before[i] != after[i] && cnt[i] == 6) ? 1 : 0;
}

__global__ void k_edit_write(const double* orig, const double* before, const int* pos,
                             const int* flag, long long n, double xi, double* edits,
                             unsigned long long* mant, int* expo) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    // cost(): edits.push_back(-(decp_data_copy[i] - (input_data[i] - bound)))
    double e = -(before[i] - (orig[i] - xi));
    int p = pos[i];
    edits[p] = e;
    unsigned long long bits = (unsigned long long)__double_as_longlong(e);
    mant[p] = bits & 0xFFFFFFFFFFFFFull;
    expo[p] = (int)((bits >> 52) & 0x7FF) - 1023;
  }
}

// Exclusive scan of flags in blocks of 1024 (block sums written out).
__global__ void __launch_bounds__(1024) k_scan_block(const int* in, int* out, long long n,
                                                     int* block_sums) {
  __shared__ int sh[1024];
  long long i = blockIdx.x * 1024LL + threadIdx.x;
  int v = i < n ? in[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    int t = threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == 1023) block_sums[blockIdx.x] = sh[1023];
}
__global__ void __launch_bounds__(1024) k_scan_add(int* out, long long n, const int* offs) {
  long long i = blockIdx.x * 1024LL + threadIdx.x;
  if (i < n) out[i] += offs[blockIdx.x];
}

}  // namespace ez

namespace ez {

// ------------------------------------------------- split tree (check) ----
// compute_MergeT (src/ExtreM.cpp:2930-2952) on the GPU.  Maxima are named by
// their position p in MX; mkey[p] is the ordered key of their
// value, so (mkey, MX) is the (value, id) order of the reference's ranks.

__device__ __forceinline__ bool max_less(const long long* mkey, const int* MX, int a, int b) {
  return mkey[a] < mkey[b] || (mkey[a] == mkey[b] && MX[a] < MX[b]);
}

// findAscPaths (src/ExtreM.cpp:1597-1636): per 2-saddle, the maxima reached
// by steepest ascent from its upper neighbours, sorted by (value, id) and
// deduplicated, kNbr slots per saddle.
__global__ void __launch_bounds__(256) k_tree_triplets(Geo g, const double* f, const int* S2,
                                                       int n2, const int* ds, const int* ext_id,
                                                       const long long* mkey, const int* MX,
                                                       int* T, int* cnt) {
  GRID_STRIDE(i, n2) {
    const int v = S2[i];
    int x, y, z;
    coords(g, v, x, y, z);
    const int key = c_cls_key[class_of(g, x, y, z)];
    const double fv = f[v];
    int loc[kNbr];
    int m = 0;
    for (int d = 0; d < kNbr; d++) {
      if (!((key >> d) & 1)) continue;
      const int r = v + g.off[d];
      const double fr = f[r];
      if (!(fr > fv || (fr == fv && r > v))) continue;
      const int p = ext_id[ds[r]] >> 1;
      int k = 0;
      bool dup = false;
      for (; k < m; k++) {
        if (loc[k] == p) { dup = true; break; }
        if (max_less(mkey, MX, p, loc[k])) break;
      }
      if (dup) continue;
      for (int j = m; j > k; j--) loc[j] = loc[j - 1];
      loc[k] = p;
      m++;
    }
    for (int k = 0; k < m; k++) T[i * kNbr + k] = loc[k];
    cnt[i] = m;
  }
}

// constructPersistencePairs, largest saddle per maximum (src/ExtreM.cpp:2146-2166):
// pass 0 raises the value key, pass 1, among the saddles with that key, the
// vertex id (high 32 bits) carrying the list position (low 32 bits).  The
// global maximum (position gmax) takes no part.
__global__ void k_tree_best(const int* T, const int* cnt, const int* S2, int n2,
                            const long long* skey, int gmax, int pass, long long* bestkey,
                            long long* bestpos) {
  GRID_STRIDE(i, n2) {
    const int m = cnt[i];
    for (int k = 0; k < m; k++) {
      const int p = T[i * kNbr + k];
      if (p == gmax) continue;
      if (pass == 0) atomicMax(&bestkey[p], skey[i]);
      else if (skey[i] == bestkey[p])
        atomicMax(&bestpos[p], ((long long)S2[i] << 32) | (unsigned)i);
    }
  }
}

// A maximum that is the lowest maximum of its largest saddle is paired with
// it and points to the saddle's highest maximum (src/ExtreM.cpp:2170-2195).
__global__ void k_tree_pair(const int* T, const int* cnt, int nmax, const long long* bestpos,
                            int* pair_pos, int* parent, int* ptr, int* changed) {
  GRID_STRIDE(p, nmax) {
    if (bestpos[p] < 0) continue;
    const int i = (int)(bestpos[p] & 0xffffffffLL);
    if (T[i * kNbr] != p) continue;
    const int hi = T[i * kNbr + cnt[i] - 1];
    pair_pos[p] = i;
    parent[p] = hi;
    ptr[p] = hi;
    *changed = 1;
  }
}

// Path compression of the maximum pointers (src/ExtreM.cpp:2197-2211):
// pointers only lead to higher maxima, so every walk ends.
__global__ void k_tree_compress(int n, int* ptr) {
  GRID_STRIDE(p, n) {
    int r = ptr[p];
    int rr = __hip_atomic_load(&ptr[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int step = 0; step < n && rr != r; step++) {
      r = rr;
      rr = __hip_atomic_load(&ptr[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    ptr[p] = r;
  }
}

// Relabels every triplet through the pointers, sorts and deduplicates it; a
// triplet left with one maximum is emptied (src/ExtreM.cpp:2213-2222).
__global__ void k_tree_relabel(int* T, int* cnt, int n2, const int* ptr, const long long* mkey,
                               const int* MX) {
  GRID_STRIDE(i, n2) {
    const int m0 = cnt[i];
    if (m0 == 0) continue;
    int loc[kNbr];
    int m = 0;
    for (int j = 0; j < m0; j++) {
      const int p = ptr[T[i * kNbr + j]];
      int k = 0;
      bool dup = false;
      for (; k < m; k++) {
        if (loc[k] == p) { dup = true; break; }
        if (max_less(mkey, MX, p, loc[k])) break;
      }
      if (dup) continue;
      for (int q = m; q > k; q--) loc[q] = loc[q - 1];
      loc[k] = p;
      m++;
    }
    if (m == 1) m = 0;
    for (int k = 0; k < m; k++) T[i * kNbr + k] = loc[k];
    cnt[i] = m;
  }
}

// Per-position value keys of a vertex list.
__global__ void k_list_keys(const int* list, int n, const double* f, long long* key) {
  GRID_STRIDE(i, n) key[i] = okey_c(f[list[i]]);
}

}  // namespace ez
