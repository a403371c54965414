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
__constant__ int c_cls_key[64];
__constant__ int c_lut_off[64];

void upload_constants(const int* dx, const int* dy, const int* dz,
                      const int* cls_key, const int* lut_off) {
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dx), dx, sizeof(int) * kNbr));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dy), dy, sizeof(int) * kNbr));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dz), dz, sizeof(int) * kNbr));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_cls_key), cls_key, sizeof(int) * 64));
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_lut_off), lut_off, sizeof(int) * 64));
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
  return axis_state(x, g.W) + 4 * axis_state(y, g.H) + 16 * axis_state(z, g.D);
}

// Appends v to list when pred; all lanes of the wave must call it.
__device__ __forceinline__ void wave_append(int* list, int* counter, bool pred, int v) {
  unsigned long long m = __ballot(pred);
  if (m == 0) return;
  int lane = __lane_id();
  int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
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
  if (lc == k) return 4;   // no upper neighbour
  if (lc == 0) return 0;   // no lower neighbour
  return lut[c_lut_off[cls] + cfg];
}

// Type of a known lower mask (LUT lookup only).
__device__ __forceinline__ int type_of_mask(const uint8_t* lut, int cls, unsigned lm) {
  int key = c_cls_key[cls];
  unsigned cfg = 0;
  int k = 0, lc = 0;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (!((key >> d) & 1)) continue;
    if ((lm >> d) & 1) lc++;
    else cfg |= 1u << k;
    k++;
  }
  if (lc == k) return 4;
  if (lc == 0) return 0;
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

// swap + the caller's `d > oldValue` gate, sequentially a max (D5).  The
// first proposal above the sentinel lists the vertex as touched.
__device__ __forceinline__ void propose(const St& s, int t, double d) {
  long long k = okey(d);
  if (k <= s.sent) return;
  long long old = atomicMax(&s.delta[t], k);
  if (old == s.sent) {
    int i = atomicAdd(&s.ctr[C_TOUCHED], 1);
    s.touched[i] = t;
  }
}
__device__ __forceinline__ double lower_mid(const St& s, int v) {
  return ((s.orig[v] - s.xi) + s.decp[v]) / 2.0 - s.decp[v];
}
__device__ __forceinline__ double raise_mid(const St& s, int v) {
  return ((s.orig[v] + s.xi) + s.decp[v]) / 2.0 - s.decp[v];
}

#define GRID_STRIDE(i, n) \
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

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
__global__ void k_f32_to_f64(const float* in, double* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = (double)in[i];
}
// Sets ctr[C_FAIL] when any value is not finite.
__global__ void k_check_finite(const double* a, long long n, int* ctr) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    if (!isfinite(a[i])) ctr[C_FAIL] = 1;
}

// ---------------------------------------------------- classification ----
// Full-grid classification of a field: block (64, 4), grid (ceil(W/64),
// ceil(H/4), D).  classifyAllVertices, src/ExtreM.cpp:1088-1097.
__global__ void __launch_bounds__(256) k_classify(Geo g, const uint8_t* lut, const double* f,
                                                  uint8_t* type, uint16_t* lower) {
  int x = blockIdx.x * 64 + threadIdx.x;
  int y = blockIdx.y * 4 + threadIdx.y;
  int z = blockIdx.z;
  if (x >= g.W || y >= g.H) return;
  int v = x + g.W * (y + g.H * z);
  unsigned lm;
  int t = classify_at(g, lut, f, v, class_of(g, x, y, z), &lm);
  type[v] = (uint8_t)t;
  lower[v] = (uint16_t)lm;
}

// Full-grid false-critical-point scan (the first get_false_criticle_points
// of c_loops, src/ExtreM.cpp:1513-1519); appends to the frontier.
__global__ void __launch_bounds__(256) k_check_full(St s, int* front, int fctr, int epoch) {
  int x = blockIdx.x * 64 + threadIdx.x;
  int y = blockIdx.y * 4 + threadIdx.y;
  int z = blockIdx.z;
  bool in = x < s.g.W && y < s.g.H;
  int v = in ? x + s.g.W * (y + s.g.H * z) : 0;
  bool mis = false;
  if (in) {
    int cls = class_of(s.g, x, y, z);
    unsigned dl = lower_mask_at(s.g, s.decp, v, cls);
    unsigned ol = s.olow[v];
    if (dl != ol) {
      int ot = s.otype[v];
      int dt = type_of_mask(s.lut, cls, dl);
      mis = dt != ot || is_saddle_type(ot);
    }
  }
  (void)epoch;
  wave_append(front, &s.ctr[fctr], mis, v);
}

// Re-check after edits: every list entry (and, with nbrs, its 14 neighbours)
// is evaluated once per epoch; mismatches go to the frontier.  Statuses of
// vertices outside the edited neighbourhoods cannot change, so old frontier
// entries plus edited neighbourhoods give the full get_false_criticle_points
// result.
__global__ void __launch_bounds__(256) k_check_list(St s, const int* list, int lctr,
                                                    int with_nbrs, int* front, int fctr,
                                                    int epoch) {
  int n = s.ctr[lctr];
  GRID_STRIDE(i, n) {
    int v = list[i];
    int cand[kNbr + 1];
    int m = 0;
    cand[m++] = v;
    if (with_nbrs) {
      int x, y, z;
      coords(s.g, v, x, y, z);
      int key = c_cls_key[class_of(s.g, x, y, z)];
      for (int d = 0; d < kNbr; d++)
        if ((key >> d) & 1) cand[m++] = v + s.g.off[d];
    }
    for (int j = 0; j < m; j++) {
      int u = cand[j];
      if (atomicExch(&s.stamp[u], epoch) == epoch) continue;
      if (mismatch_at(s, u)) {
        int p = atomicAdd(&s.ctr[fctr], 1);
        front[p] = u;
      }
    }
  }
}

// Clears the dirty flags of the listed vertices and the list counter.
__global__ void k_clear_dirty(St s) {
  int n = s.ctr[C_DIRTY];
  GRID_STRIDE(i, n) s.dflag[s.dirty[i]] = 0;
}

// ------------------------------------------------- c_loops proposals ----
// c_loop (src/ExtreM.cpp:1152-1236) for false maxima / minima and
// get_false_stars (src/ExtreM.cpp:1412-1445) for saddle-list entries.
__global__ void __launch_bounds__(256) k_propose_c(St s, const int* front, int fctr) {
  int n = s.ctr[fctr];
  GRID_STRIDE(i, n) {
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
        propose(s, v, lower_mid(s, v));
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
        if (!(fv > fl || (fv == fl && v > largest))) propose(s, largest, lower_mid(s, largest));
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
          propose(s, smallest, lower_mid(s, smallest));
      } else {
        propose(s, v, lower_mid(s, v));
      }
    }
    bool sad = (dt != ot) ? (is_saddle_type(dt) || is_saddle_type(ot))
                          : (is_saddle_type(ot) && dl != ol);
    if (sad && dl != ol) {  // get_false_stars
      unsigned a = ol & ~dl;
      for (int d = 0; d < kNbr; d++)
        if ((a >> d) & 1) {
          int r = v + s.g.off[d];
          propose(s, r, (s.orig[r] - s.xi) - s.decp[r]);
        }
      if (dl & ~ol) propose(s, v, (s.orig[v] - s.xi) - s.decp[v]);
    }
  }
}

// apply_delta, split (src/ExtreM.cpp:1486-1503), over the touched list only;
// resets the delta slot and records changed vertices as dirty.
__global__ void __launch_bounds__(256) k_apply(St s) {
  int n = s.ctr[C_TOUCHED];
  GRID_STRIDE(i, n) {
    int t = s.touched[i];
    double dl = okey_inv(s.delta[t]);
    s.delta[t] = s.sent;
    double o = s.orig[t], x = s.decp[t];
    int c = s.cnt[t];
    double nx;
    if (fabs(dl) > 1e-15 && c < 5 && fabs(o - x - dl) <= s.xi) {
      nx = x + dl;
      c += 1;
    } else {
      nx = o - s.xi;
      c = 6;
    }
    s.decp[t] = nx;
    s.cnt[t] = (uint8_t)c;
    if (nx != x) {
      s.ctr[C_ANY_CHANGE] = 1;
      if (!s.dflag[t]) {
        s.dflag[t] = 1;
        int p = atomicAdd(&s.ctr[C_DIRTY], 1);
        s.dirty[p] = t;
      }
    }
  }
}

// ----------------------------------------------------- manifold labels ----
// Steepest ascent (ds) and descent (as) pointer of every vertex, first loops
// of src/ExtreM.cpp:146-157 and 504-515.
__global__ void __launch_bounds__(256) k_steepest(Geo g, const double* f, int* ds, int* as) {
  int x = blockIdx.x * 64 + threadIdx.x;
  int y = blockIdx.y * 4 + threadIdx.y;
  int z = blockIdx.z;
  if (x >= g.W || y >= g.H) return;
  int v = x + g.W * (y + g.H * z);
  int key = c_cls_key[class_of(g, x, y, z)];
  double c = f[v];
  int up = v, dn = v;
  double fu = c, fd = c;
#pragma unroll
  for (int d = 0; d < kNbr; d++) {
    if (!((key >> d) & 1)) continue;
    int r = v + g.off[d];
    double fr = f[r];
    if (fu < fr || (fu == fr && up < r)) {
      up = r;
      fu = fr;
    }
    if (fd > fr || (fd == fr && dn > r)) {
      dn = r;
      fd = fr;
    }
  }
  ds[v] = up;
  as[v] = dn;
}

// One in-place pointer-jumping round on both label arrays (D7: every vertex
// is compressed to its root).  Sets ctr[C_CHANGED] when anything moved.
__global__ void __launch_bounds__(256) k_jump(int n, int* ds, int* as, int* ctr) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  bool ch = false;
  if (v < n) {
    int r = ds[v], rr = ds[r];
    if (r != rr) {
      ds[v] = rr;
      ch = true;
    }
    r = as[v];
    rr = as[r];
    if (r != rr) {
      as[v] = rr;
      ch = true;
    }
  }
  if (__ballot(ch) && __lane_id() == 0) ctr[C_CHANGED] = 1;
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

// The extremum of a list with the highest (want_max) or lowest (value, id).
__global__ void __launch_bounds__(256) k_list_best(const int* list, int n, const double* f,
                                                   int* slot, int want_max) {
  int best = kNone;
  double fb = 0.0;
  GRID_STRIDE(i, n) {
    int v = list[i];
    double fv = f[v];
    if (best == kNone || (want_max ? larger(fv, v, fb, best) : smaller(fv, v, fb, best))) {
      best = v;
      fb = fv;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    int ob = __shfl_xor(best, o);
    double of = __shfl_xor(fb, o);
    if (ob != kNone &&
        (best == kNone || (want_max ? larger(of, ob, fb, best) : smaller(of, ob, fb, best)))) {
      best = ob;
      fb = of;
    }
  }
  if (__lane_id() == 0 && best != kNone) cas_best(slot, best, f, want_max != 0);
}

__global__ void k_list_fill(const int* list, int n, int* arr, int val) {
  GRID_STRIDE(i, n) arr[list[i]] = val;
}

// computeCP + findAscPaths/findDesPaths + computelargestSaddlesForMax /
// computesmallestSaddlesForMin for one saddle list (src/ExtreM.cpp:1597-1679,
// 1920-2040, 2309-2371), without sorting:
//  side 2: upper neighbours, ascent roots L; first = lowest (value,id) root,
//          last = highest; each root that is a listed maximum other than the
//          global one keeps the highest (value,id) saddle reaching it.
//  side 1: lower neighbours, descent roots L; first = highest root, last =
//          lowest; each listed minimum other than the global one keeps the
//          lowest saddle.
// With ormap, also stores the roots in neighbour order (compute_Max_for_Saddle,
// src/ExtreM.cpp:2373-2403; 14 slots, D3).
__global__ void __launch_bounds__(256) k_saddles(St s, int side, const int* list, int n,
                                                 const double* f, const int* L, int* first,
                                                 int* last, int* slots, int glob_idx,
                                                 int* ormap) {
  int glob = s.ctr[glob_idx];
  int ext_type = side == 2 ? 4 : 0;
  GRID_STRIDE(i, n) {
    int sv = list[i];
    int x, y, z;
    coords(s.g, sv, x, y, z);
    int key = c_cls_key[class_of(s.g, x, y, z)];
    double fs = f[sv];
    int lo = kNone, hi = kNone, c = 0;
    double flo = 0.0, fhi = 0.0;
    int roots[kNbr];
    for (int d = 0; d < kNbr; d++) {
      if (!((key >> d) & 1)) continue;
      int r = sv + s.g.off[d];
      double fr = f[r];
      bool take = side == 2 ? larger(fr, r, fs, sv) : smaller(fr, r, fs, sv);
      if (!take) continue;
      int root = L[r];
      double fo = f[root];
      roots[c++] = root;
      if (lo == kNone || smaller(fo, root, flo, lo)) {
        lo = root;
        flo = fo;
      }
      if (hi == kNone || larger(fo, root, fhi, hi)) {
        hi = root;
        fhi = fo;
      }
    }
    if (side == 2) {
      first[i] = lo;
      last[i] = hi;
    } else {
      first[i] = hi;
      last[i] = lo;
    }
    if (ormap) {
      for (int k = 0; k < kNbr; k++) ormap[(long long)i * kNbr + k] = k < c ? roots[k] : kNone;
    }
    for (int k = 0; k < c; k++) {
      int r = roots[k];
      if (r == glob || s.otype[r] != ext_type) continue;
      bool dup = false;
      for (int j = 0; j < k; j++) dup |= roots[j] == r;
      if (dup) continue;
      cas_best(&slots[r], sv, f, side == 2);
    }
  }
}

// ------------------------------------------- wrong neighbours (greedy) ----
// Candidate masks of get_wrong_split_neighbors / get_wrong_join_neighbors
// (src/ExtreM.cpp:2405-2450): bit d set when the d-th neighbour is in the
// saddle's upper (lower) star of decp and its root differs from the root of
// the same-rank original neighbour.  Entries e < n2 are 2-saddles, the rest
// 1-saddles.
__global__ void __launch_bounds__(256) k_wn_cand(St s, int n2, int n1, int* active, int actr) {
  int n = n2 + n1;
  GRID_STRIDE(e, n) {
    int side2 = e < n2;
    int i = side2 ? e : e - n2;
    int sv = side2 ? s.S2[i] : s.S1[i];
    const int* L = side2 ? s.dDS : s.dAS;
    const int* om = side2 ? s.ormap2 : s.ormap1;
    int x, y, z;
    coords(s.g, sv, x, y, z);
    int key = c_cls_key[class_of(s.g, x, y, z)];
    double fs = s.decp[sv];
    unsigned mask = 0;
    int c = 0;
    for (int d = 0; d < kNbr; d++) {
      if (!((key >> d) & 1)) continue;
      int r = sv + s.g.off[d];
      double fr = s.decp[r];
      bool take = side2 ? larger(fr, r, fs, sv) : smaller(fr, r, fs, sv);
      if (!take) continue;
      int orl = c < kNbr ? om[(long long)i * kNbr + c] : kNone;
      if (L[r] != orl) mask |= 1u << d;
      c++;
    }
    s.cmask[e] = (uint16_t)mask;
    if (mask) {
      int p = atomicAdd(&s.ctr[actr], 1);
      active[p] = e;
    }
  }
}

// Deterministic reservations reproducing the serial greedy in saddle-id
// order: drop flagged candidates, then every remaining candidate records the
// smallest saddle id that may still take it.
__global__ void __launch_bounds__(256) k_wn_register(St s, int n2, const int* active, int actr) {
  int n = s.ctr[actr];
  GRID_STRIDE(a, n) {
    int e = active[a];
    int side2 = e < n2;
    int sv = side2 ? s.S2[e] : s.S1[e - n2];
    const uint8_t* fl = side2 ? s.wn2 : s.wn1;
    int* ms = side2 ? s.minsad2 : s.minsad1;
    unsigned mask = s.cmask[e];
    for (int d = 0; d < kNbr; d++) {
      if (!((mask >> d) & 1)) continue;
      int r = sv + s.g.off[d];
      if (fl[r]) {
        mask &= ~(1u << d);
        continue;
      }
      atomicMin(&ms[r], sv);
    }
    s.cmask[e] = (uint16_t)mask;
  }
}

// A saddle whose first candidate is reserved by itself takes it (flag +
// append, as get_wrong_split_neighbors does before its return).
__global__ void __launch_bounds__(256) k_wn_commit(St s, int n2, const int* active, int actr) {
  int n = s.ctr[actr];
  GRID_STRIDE(a, n) {
    int e = active[a];
    unsigned mask = s.cmask[e];
    if (!mask) continue;
    int side2 = e < n2;
    int sv = side2 ? s.S2[e] : s.S1[e - n2];
    int d = __ffs(mask) - 1;
    int r = sv + s.g.off[d];
    int* ms = side2 ? s.minsad2 : s.minsad1;
    if (ms[r] == sv) {
      if (side2) {
        s.wn2[r] = 1;
        int p = atomicAdd(&s.ctr[C_NFC], 1);
        s.wnl2[p] = r;
      } else {
        s.wn1[r] = 1;
        int p = atomicAdd(&s.ctr[C_NFC1], 1);
        s.wnl1[p] = r;
      }
      s.cmask[e] = (uint16_t)(mask | 0x8000u);  // done
    }
  }
}

// Releases this round's reservations and keeps unfinished saddles.
__global__ void __launch_bounds__(256) k_wn_next(St s, int n2, const int* active, int actr,
                                                 int* next, int nctr) {
  int n = s.ctr[actr];
  GRID_STRIDE(a, n) {
    int e = active[a];
    unsigned mask = s.cmask[e];
    int side2 = e < n2;
    int sv = side2 ? s.S2[e] : s.S1[e - n2];
    int* ms = side2 ? s.minsad2 : s.minsad1;
    for (int d = 0; d < kNbr; d++)
      if ((mask >> d) & 1) ms[sv + s.g.off[d]] = 0x7fffffff;
    if (mask && !(mask & 0x8000u)) {
      int p = atomicAdd(&s.ctr[nctr], 1);
      next[p] = e;
    }
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
__global__ void __launch_bounds__(256) k_fixpath(St s) {
  int n2 = s.ctr[C_NFC], n1 = s.ctr[C_NFC1];
  GRID_STRIDE(e, n2 + n1) {
    int dir = e < n2 ? 0 : 1;
    int cur = dir == 0 ? s.wnl2[e] : s.wnl1[e - n2];
    int tru = kNone, fal = kNone;
    for (long long step = 0; step <= (long long)s.g.N; step++) {
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
    if (tru == fal) continue;
    if (dir == 0) propose(s, fal, lower_mid(s, fal));
    else propose(s, tru, lower_mid(s, tru));
  }
}

// ------------------------------------------------- wrong index (ranks) ----
// get_wrong_index_max / get_wrong_index_min (src/ExtreM.cpp:2595-2670): the
// first saddle (lowest id) per original extremum wins; pairs are compared by
// vertex id.  phase 0 registers, phase 1 counts (and proposes, s_loops /
// s_loops_join, src/ExtreM.cpp:2838-2886), phase 2 releases.
__global__ void __launch_bounds__(256) k_wrong_index(St s, int side, int phase, int do_fix) {
  int n = side == 2 ? s.n2 : s.n1;
  const int* S = side == 2 ? s.S2 : s.S1;
  const int* of = side == 2 ? s.of2 : s.of1;
  const int* ol = side == 2 ? s.ol2 : s.ol1;
  const int* df = side == 2 ? s.df2 : s.df1;
  const int* dl = side == 2 ? s.dl2 : s.dl1;
  GRID_STRIDE(i, n) {
    int sv = S[i];
    bool mf = of[i] != df[i], ml = ol[i] != dl[i];
    if (phase == 0) {
      if (mf) atomicMin(&s.win_a[of[i]], sv);
      if (ml) atomicMin(&s.win_b[ol[i]], sv);
    } else if (phase == 1) {
      if (mf && s.win_a[of[i]] == sv) {
        if (side == 2) {  // [0] -> wrong_max_counter, fix dir 0 lowers orig
          atomicAdd(&s.ctr[C_WMC], 1);
          if (do_fix) propose(s, of[i], lower_mid(s, of[i]));
        } else {          // [0] -> wrong_min_counter_2, fix dir 1 lowers dec
          atomicAdd(&s.ctr[C_WMINC2], 1);
          if (do_fix && df[i] != kNone) propose(s, df[i], lower_mid(s, df[i]));
        }
      }
      if (ml && s.win_b[ol[i]] == sv) {
        if (side == 2) {  // [last] -> wrong_max_counter_2, fix dir 1 lowers dec
          atomicAdd(&s.ctr[C_WMC2], 1);
          if (do_fix && dl[i] != kNone) propose(s, dl[i], lower_mid(s, dl[i]));
        } else {          // [last] -> wrong_min_counter, fix dir 0 lowers orig
          atomicAdd(&s.ctr[C_WMINC], 1);
          if (do_fix) propose(s, ol[i], lower_mid(s, ol[i]));
        }
      }
    } else {
      if (mf) s.win_a[of[i]] = 0x7fffffff;
      if (ml) s.win_b[ol[i]] = 0x7fffffff;
    }
  }
}

// get_wrong_index_saddles(_join) + fix_wrong_index_saddle
// (src/ExtreM.cpp:2672-2715, 2811-2836).  A side without an adjacent saddle
// (the reference reads one past its saddle vector there) is no mismatch.
__global__ void __launch_bounds__(256) k_wrong_saddle(St s, int side, int do_fix) {
  int n = side == 2 ? s.nmax : s.nmin;
  const int* E = side == 2 ? s.MX : s.MN;
  const int* os = side == 2 ? s.olarg : s.osmall;
  const int* ds = side == 2 ? s.dlarg : s.dsmall;
  GRID_STRIDE(i, n) {
    int m = E[i];
    int ov = os[m], dv = ds[m];
    if (ov == dv || ov == kNone || dv == kNone) continue;  // D12
    atomicAdd(&s.ctr[side == 2 ? C_WSC : C_WSCJ], 1);
    if (!do_fix) continue;
    double ft = s.decp[ov], ff = s.decp[dv];
    if (ft > ff || (ft == ff && ov > dv)) continue;
    if (side == 2) propose(s, dv, lower_mid(s, dv));
    else propose(s, ov, raise_mid(s, ov));
  }
}

// --------------------------------------------------------------- setup ----
// Appends every vertex to the saddle / extremum lists by original type.
__global__ void __launch_bounds__(256) k_lists(St s) {
  int v = blockIdx.x * blockDim.x + threadIdx.x;
  bool in = v < s.g.N;
  int t = in ? s.otype[v] : 5;
  wave_append(s.S2, &s.ctr[C_N2], in && (t == 2 || t == 3), v);
  wave_append(s.S1, &s.ctr[C_N1], in && (t == 1 || t == 3), v);
  wave_append(s.MX, &s.ctr[C_NMAX], in && t == 4, v);
  wave_append(s.MN, &s.ctr[C_NMIN], in && t == 0, v);
}

// main.cpp:74-75 and 122-131: decp := orig - xi at the global min / max and
// at every original 1- and 2-saddle.
__global__ void k_presnap(St s) {
  int n = s.n2 + s.n1 + 2;
  GRID_STRIDE(e, n) {
    int v;
    if (e == 0) v = s.ctr[C_GMIN];
    else if (e == 1) v = s.ctr[C_GMAX];
    else if (e - 2 < s.n2) v = s.S2[e - 2];
    else v = s.S1[e - 2 - s.n2];
    s.decp[v] = s.orig[v] - s.xi;
  }
}

__global__ void k_argext(const double* f, int n, int* slot_min, int* slot_max) {
  int bmin = kNone, bmax = kNone;
  double fmin = 0.0, fmax = 0.0;
  GRID_STRIDE(v, n) {
    double fv = f[v];
    if (bmin == kNone || smaller(fv, v, fmin, bmin)) { bmin = v; fmin = fv; }
    if (bmax == kNone || larger(fv, v, fmax, bmax)) { bmax = v; fmax = fv; }
  }
  for (int o = 32; o > 0; o >>= 1) {
    int b = __shfl_xor(bmin, o);
    double fb = __shfl_xor(fmin, o);
    if (b != kNone && (bmin == kNone || smaller(fb, b, fmin, bmin))) { bmin = b; fmin = fb; }
    b = __shfl_xor(bmax, o);
    fb = __shfl_xor(fmax, o);
    if (b != kNone && (bmax == kNone || larger(fb, b, fmax, bmax))) { bmax = b; fmax = fb; }
  }
  if (__lane_id() == 0 && bmin != kNone) {
    cas_best(slot_min, bmin, f, false);
    cas_best(slot_max, bmax, f, true);
  }
}

// Number of vertices with a non-zero edit count.
__global__ void k_count_edited(const uint8_t* cnt, long long n, int* ctr) {
  int c = 0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    c += cnt[i] != 0;
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0 && c) atomicAdd(&ctr[C_EDITED], c);
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
