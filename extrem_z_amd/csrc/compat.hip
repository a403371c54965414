// compat.hip -- libExtreMz: the reference library's API (include/api/ExtreM.h)
// over the MI355X engine.
//
// The reference's main.cpp drives its correction through ~30 functions that
// read and write ~60 globals (include/api/ExtreM.h:17-124, main.cpp:65-507).
// This file defines those globals and functions:
//  - single passes over one field (classification, manifolds, sorting, CP
//    lists and triplets, the split tree) run on the GPU on an auxiliary
//    context and write the caller's arrays and containers;
//  - the correction steps (c_loops, r_loops, s_loops, saddle_loops, the
//    wrong-index and wrong-neighbour functions) run on one device-resident
//    correction (ez::Stepper) that follows the caller's calls step by step,
//    so the sequence of main.cpp gives extremz_correct's field bit for bit;
//  - the per-saddle functions the caller loops over (compute_Max_for_Saddle,
//    get_wrong_split_neighbors, ...) read results the GPU computed once per
//    loop (labels, wrong-neighbour candidates) and do the caller-visible
//    serial step (the slot write, the greedy flag-and-append) on the host.
// Host-side loops here are over containers the reference API hands over
// (the CP vectors, maps, a triplet): no field pass runs on the CPU.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/api/ExtreM.h"
#include "engine.h"

// ----------------------------------------------------------- globals ----
// include/api/ExtreM.h:18-41 (definitions src/ExtreM.cpp:15-70).
extern "C" {
// the 14 stencil directions (x, y, z), also the engine's order
int directions[42] = {1, 0, 0,  -1, 0, 0,  0, 1, 0,  0, -1, 0,  0, 0, 1,  0, 0, -1,
                      -1, 1, 0, 1, -1, 0,  0, 1, 1,  0, -1, -1, -1, 0, 1,  1, 0, -1,
                      -1, 1, 1, 1, -1, -1};
// offsets of the 26 neighbours in the reference's order (src/ExtreM.cpp:56-70)
int direction_to_index_mapping[26][3] = {
    {1, 0, 0},  {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {-1, 1, 0}, {1, -1, 0},  {0, 0, -1},
    {0, -1, 1}, {0, 0, 1},   {0, 1, -1},  {-1, 0, 1},  {1, 0, -1}, {1, 1, 0},   {-1, -1, 0},
    {1, 0, 1},  {-1, 0, -1}, {0, 1, 1},   {0, -1, -1}, {1, 1, 1},  {1, 1, -1},  {1, -1, 1},
    {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
double* d_deltaBuffer = nullptr;
int width = 0, height = 0, depth = 0, maxNeighbors = 0, num_Elements = 0;
int number_of_false_cases = 0, number_of_false_cases1 = 0;
std::atomic_int count_f_max{0}, count_f_min{0}, count_f_saddle{0};
int wrong_max_counter = 0, globalMin = 0, dec_globalMin = 0, wrong_max_counter_2 = 0;
int wrong_saddle_counter = 0, wrong_saddle_counter_join = 0;
int wrong_min_counter = 0, wrong_min_counter_2 = 0;
int *adjacency = nullptr, *dec_vertex_type = nullptr, *vertex_type = nullptr;
int *all_max = nullptr, *all_min = nullptr, *all_saddle = nullptr;
double *decp_data = nullptr, *input_data = nullptr, *decp_data_copy = nullptr;
double bound = 0.0;
int *or_saddle_max_map = nullptr, *wrong_neighbors = nullptr, *wrong_neighbors1 = nullptr;
int *wrong_neighbors_index = nullptr, *wrong_rank_max = nullptr, *wrong_rank_max_index = nullptr;
int *wrong_rank_saddle = nullptr, *wrong_rank_saddle_index = nullptr;
int *wrong_rank_max_2 = nullptr, *wrong_rank_max_index_2 = nullptr;
int *wrong_rank_ds_saddle = nullptr, *wrong_rank_ds_saddle_index = nullptr;
int *or_saddle_min_map = nullptr, *wrong_neighbors_ds = nullptr, *wrong_neighbors_ds_index = nullptr;
int *wrong_rank_min = nullptr, *wrong_rank_min_index = nullptr;
int *wrong_rank_min_2 = nullptr, *wrong_rank_min_index_2 = nullptr;
int *wrong_rank_saddle_join = nullptr, *wrong_rank_saddle_join_index = nullptr;
std::vector<std::array<int, 46>> saddleTriplets, dec_saddleTriplets;
std::vector<std::array<int, 46>> saddle1Triplets, dec_saddle1Triplets;
std::vector<std::vector<int>> vertex_cells;
std::vector<int> saddles2, maximum, dec_saddles2, dec_maximum, delta_counter;
std::vector<int> saddles1, minimum, dec_saddles1, dec_minimum;
std::unordered_map<int, int> largestSaddlesForMax, dec_largestSaddlesForMax;
std::unordered_map<int, int> smallestSaddlesForMin, dec_smallestSaddlesForMin;
int *lowerStars = nullptr, *upperStars = nullptr, *dec_lowerStars = nullptr,
    *dec_upperStars = nullptr;
}

namespace ezc {

using ez::hip_check;

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error("libExtreMz: " + m); }

// Grow-only device buffer, held until the process ends (no destructor: the
// HIP runtime may be gone when statics are destroyed).
template <class T>
struct Dev {
  T* p = nullptr;
  size_t n = 0;
  T* ensure(size_t k) {
    if (k <= n && p) return p;
    if (p) hip_check(hipFree(p), "hipFree");
    p = nullptr;
    n = 0;
    if (hipMalloc((void**)&p, std::max<size_t>(k, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      throw std::bad_alloc();
    }
    n = std::max<size_t>(k, 1);
    return p;
  }
};

__constant__ int c_dir[42];

struct Grid {
  int W, H, D;
  long long N;
};

__device__ __forceinline__ int nbr(const Grid& g, int v, int d) {
  const int x = v % g.W, y = (v / g.W) % g.H, z = v / (g.W * g.H);
  const int nx = x + c_dir[3 * d], ny = y + c_dir[3 * d + 1], nz = z + c_dir[3 * d + 2];
  if (nx < 0 || nx >= g.W || ny < 0 || ny >= g.H || nz < 0 || nz >= g.D) return -1;
  return nx + g.W * (ny + g.H * nz);
}

// computeAdjacency (src/ExtreM.cpp:543-575): the in-bounds neighbours among
// the first M table directions, packed to the front, -1 after.
__global__ void k_adjacency(Grid g, int M, int* adj) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N;
       v += (long long)gridDim.x * blockDim.x) {
    int k = 0;
    for (int d = 0; d < M; d++) {
      const int r = nbr(g, (int)v, d);
      if (r >= 0) adj[v * M + k++] = r;
    }
    for (; k < M; k++) adj[v * M + k] = -1;
  }
}

// The star lists of classifyVertex1 (src/ExtreM.cpp:723-748) over the first
// M directions: lower / upper neighbours in adjacency order at
// [v*(M+1), v*(M+1)+count), -1 up to M, the count at [v*(M+1)+M].
__global__ void k_stars(Grid g, int M, const double* f, int* lower, int* upper) {
  for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N;
       v += (long long)gridDim.x * blockDim.x) {
    const double fv = f[v];
    int nl = 0, nu = 0;
    int* lo = lower ? lower + v * (M + 1) : nullptr;
    int* up = upper ? upper + v * (M + 1) : nullptr;
    for (int d = 0; d < M; d++) {
      const int r = nbr(g, (int)v, d);
      if (r < 0) continue;
      const double fr = f[r];
      if (fr < fv || (fr == fv && r < v)) {
        if (lo) lo[nl] = r;
        nl++;
      } else if (fr > fv || (fr == fv && r > v)) {
        if (up) up[nu] = r;
        nu++;
      }
    }
    if (lo) {
      for (int k = nl; k < M; k++) lo[k] = -1;
      lo[M] = nl;
    }
    if (up) {
      for (int k = nu; k < M; k++) up[k] = -1;
      up[M] = nu;
    }
  }
}

__global__ void k_widen(const uint8_t* in, int* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[i] = in[i];
}

// Ascending order of doubles as unsigned keys; -0.0 and +0.0 are one value
// (the reference's comparators treat them as equal).
__device__ __forceinline__ unsigned long long order_key(double x) {
  if (x == 0.0) x = 0.0;
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | (1ULL << 63));
}

// keys of the entries of list (identity when list is null)
__global__ void k_keys(const double* f, const int* list, long long n, unsigned long long* keys,
                       int* ids) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int v = list ? list[i] : (int)i;
    keys[i] = order_key(f[v]);
    ids[i] = v;
  }
}

// computeCP's type tests (src/ExtreM.cpp:1946-1966): 2-saddles (2, 3),
// 1-saddles (1, 3), maxima (4), minima (0).
__global__ void k_cp_flags(const int* t, long long n, uint8_t* f2, uint8_t* f1, uint8_t* fx,
                           uint8_t* fn) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int y = t[i];
    f2[i] = y == 2 || y == 3;
    f1[i] = y == 1 || y == 3;
    fx[i] = y == 4;
    fn[i] = y == 0;
  }
}

__global__ void k_reverse(const int* in, int* out, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    out[n - 1 - i] = in[i];
}

__global__ void k_rank(const int* list, long long n, int* rank) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    rank[list[i]] = (int)i;
}

// findAscPaths / findDesPaths + sortAndRemoveDuplicates (src/ExtreM.cpp:
// 1597-1679, 585-595) for list entry i: the ranks (rank[lab[r]]) of the
// upper (up) or lower neighbours r of the saddle over the first M
// directions, sorted, duplicates removed in place (the tail keeps the
// leftover entries), count at [44], the saddle at [45].  A label outside
// the field or a root without a rank gives -1 (the reference reads an
// uninitialised rank there).
__global__ void k_triplets(Grid g, int M, const int* L, long long n, const double* f,
                           const int* lab, const int* rank, int up, int* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int v = L[i];
    int t[44];
    int c = 0;
    const double fv = f[v];
    for (int d = 0; d < M; d++) {
      const int r = nbr(g, v, d);
      if (r < 0) continue;
      const double fr = f[r];
      const bool keep = up ? (fr > fv || (fr == fv && r > v)) : (fr < fv || (fr == fv && r < v));
      if (!keep) continue;
      const int l = lab[r];
      t[c++] = (l >= 0 && l < g.N) ? rank[l] : -1;
    }
    for (int a = 1; a < c; a++) {  // insertion sort
      const int x = t[a];
      int b = a - 1;
      while (b >= 0 && t[b] > x) {
        t[b + 1] = t[b];
        b--;
      }
      t[b + 1] = x;
    }
    int u = c > 0 ? 1 : 0;
    for (int p = 1; p < c; p++)
      if (t[p - 1] != t[p]) t[u++] = t[p];
    int* o = out + i * 46;
    for (int k = 0; k < 44; k++) o[k] = k < c ? t[k] : 0;
    // sortAndRemoveDuplicates of an empty triplet sets the count to 1
    o[44] = c > 0 ? u : 1;
    o[45] = v;
  }
}

__global__ void k_fill(int* p, int v, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_gather(const int* src, long long nsrc, const int* idx, long long n, int* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int j = idx[i];
    out[i] = (j >= 0 && j < nsrc) ? src[j] : -1;
  }
}

// getdata's lossy field: decp = 2*xi*rint(orig/(2*xi)), stepped one cell
// toward orig where that misses the bound, else orig itself.
__global__ void k_quantise(const double* o, double* d, long long n, double xi) {
  const double q = 2.0 * xi;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double ov = o[i];
    double dv = q * rint(ov / q);
    if (fabs(dv - ov) > xi) dv -= (dv > ov ? 1.0 : (dv < ov ? -1.0 : 0.0)) * q;
    if (fabs(dv - ov) > xi) dv = ov;
    d[i] = dv;
  }
}

// get_false_criticle_points (src/ExtreM.cpp:1324-1383): per vertex, the
// number of entries it gets in all_max / all_min / all_saddle.
__global__ void k_false_counts(long long n, int M, const int* t, const int* dt, const int* ls,
                               const int* us, const int* dls, const int* dus, int* cmax, int* cmin,
                               int* csad) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int a = dt[i], b = t[i];
    int nx = 0, nn = 0, ns = 0;
    if (a != b) {
      nx = (a == 4) != (b == 4);
      nn = (a == 0) != (b == 0);
      ns = ((a == 2) != (b == 2)) + ((a == 1) != (b == 1)) + ((a == 3) != (b == 3));
    } else if ((a == 1 || a == 2 || a == 3) && ls && us && dls && dus) {
      const long long o = i * (M + 1);
      bool diff = false;
      for (int pass = 0; pass < 2 && !diff; pass++) {
        const int* x = pass ? us + o : ls + o;
        const int* y = pass ? dus + o : dls + o;
        const int cx = x[M], cy = y[M];
        if (cx != cy) {
          diff = true;
          break;
        }
        for (int k = 0; k < cx && k < M; k++)
          if (x[k] != y[k]) diff = true;
      }
      ns = diff;
    }
    cmax[i] = nx;
    cmin[i] = nn;
    csad[i] = ns;
  }
}

__global__ void k_false_write(long long n, const int* c, const int* off, int* out) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    for (int k = 0; k < c[i]; k++) out[off[i] + k] = (int)i;
}

int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 4096)); }

// ------------------------------------------------------------- state ----
struct Lib {
  std::recursive_mutex mu;  // one API call at a time (the API is global state)
  extremz_ctx* work = nullptr;  // the correction
  extremz_ctx* aux = nullptr;   // single passes
  // the correction: device copies of the caller's fields
  ez::Stepper* st = nullptr;
  Dev<double> orig, decp;
  Dev<uint8_t> cnt;
  int W = 0, H = 0, D = 0, planar = 0;
  double xi = 0.0;
  const double* orig_host = nullptr;
  std::vector<double> shadow;  // decp_data as last synchronised
  std::vector<uint8_t> hcnt;
  // single-pass scratch
  Dev<double> f;
  Dev<int> ia, ib, ic, id, ie, rank, lists, sorted_ids, trip;
  Dev<unsigned long long> ka, kb;
  Dev<uint8_t> u8a, u8b, u8c, u8d, tmp;
  Dev<uint16_t> u16;
  Dev<int> nsel;
  // generation: bumped by every call except the per-saddle ones, whose
  // caches below are built once per generation
  std::atomic<unsigned long long> gen{1};
  struct LabelCache {
    std::atomic<unsigned long long> gen{0};
    const double* field = nullptr;
    std::vector<int> ds, as;
  } lab;
  struct CandCache {
    std::atomic<unsigned long long> gen{0};
    std::unordered_map<int, unsigned> split, join;  // saddle -> candidate mask
  } cand;
};

Lib& L() {
  static Lib* l = new Lib();  // never destroyed: device memory is held until exit
  return *l;
}

hipStream_t aux_stream() { return ez::ctx_stream(L().aux); }

void ensure_device() {
  Lib& l = L();
  if (l.work) return;
  l.work = ez::new_ctx(0);
  l.aux = l.work ? ez::new_ctx(0) : nullptr;
  if (!l.work || !l.aux) fail("no gfx950 GPU (HIP device 0) is available");
  hip_check(hipMemcpyToSymbol(HIP_SYMBOL(c_dir), directions, sizeof(directions)), "directions");
}

long long nvox() {
  if (!ez::valid_shape(width, height, depth)) fail("width / height / depth not set or too large");
  const long long n = (long long)width * height * depth;
  if (num_Elements != n) fail("num_Elements differs from width * height * depth");
  return n;
}

Grid grid() { return Grid{width, height, depth, nvox()}; }

// The adjacency of the correction: the reference CPU path's plane adjacency
// for maxNeighbors 6 on a depth-1 field, the 14-direction stencil otherwise.
int planar_mode() {
  if (maxNeighbors == 6 && depth == 1) return 1;
  if (maxNeighbors == 6 || maxNeighbors == 14) return 0;
  fail("maxNeighbors must be 6 (the CPU path's) or 14 (the stencil)");
}
// directions of the correction's adjacency (first table entries, in bounds)
int engine_dirs() { return planar_mode() ? 6 : 14; }

void check_split(const std::string& type) {
  if (type != "split") fail("tree type \"" + type + "\" is not supported (only \"split\")");
}

template <class T>
void h2d(T* d, const T* h, long long n, hipStream_t q) {
  if (n) hip_check(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, q), "h2d");
}
template <class T>
void d2h(T* h, const T* d, long long n, hipStream_t q) {
  if (n) hip_check(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, q), "d2h");
}
void sync(hipStream_t q) { hip_check(hipStreamSynchronize(q), "hipStreamSynchronize"); }

// uploads a caller field to the single-pass buffer
const double* upload_field(const double* h) {
  Lib& l = L();
  const long long n = nvox();
  double* d = l.f.ensure(n);
  h2d(d, h, n, aux_stream());
  return d;
}

// sorted (value, id) ascending order of the n entries of d_list (null: all
// vertices) of field d_f into d_out
void sort_by_value(const double* d_f, const int* d_list, long long n, int* d_out) {
  Lib& l = L();
  hipStream_t q = aux_stream();
  if (n == 0) return;
  unsigned long long* ka = l.ka.ensure(n);
  unsigned long long* kb = l.kb.ensure(n);
  int* ids = l.sorted_ids.ensure(n);
  k_keys<<<grid_for(n), 256, 0, q>>>(d_f, d_list, n, ka, ids);
  size_t bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, ka, kb, ids, d_out, (int)n, 0, 64,
                                               q),
            "sort size");
  uint8_t* tmp = l.tmp.ensure(bytes);
  hip_check(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, ka, kb, ids, d_out, (int)n, 0, 64, q),
            "radix sort");
}

// indices i with flag[i] != 0, ascending, into out; returns the count
long long select_flagged(const uint8_t* flags, long long n, int* out) {
  Lib& l = L();
  hipStream_t q = aux_stream();
  int* ns = l.nsel.ensure(1);
  hipcub::CountingInputIterator<int> it(0);
  size_t bytes = 0;
  hip_check(hipcub::DeviceSelect::Flagged(nullptr, bytes, it, flags, out, ns, (int)n, q),
            "select size");
  uint8_t* tmp = l.tmp.ensure(bytes);
  hip_check(hipcub::DeviceSelect::Flagged(tmp, bytes, it, flags, out, ns, (int)n, q), "select");
  int h = 0;
  d2h(&h, ns, 1, q);
  sync(q);
  return h;
}

// ------------------------------------------------------ the correction ----
bool session_current() {
  Lib& l = L();
  return l.st && l.W == width && l.H == height && l.D == depth && l.orig_host == input_data &&
         l.xi == bound && l.planar == planar_mode();
}

void session_end() {
  Lib& l = L();
  if (l.st) ez::stepper_free(l.st);
  l.st = nullptr;
  l.cand.gen.store(0);
}

// The device correction of the current field, begun on first use
// (main.cpp:139, the first c_loops): input_data and decp_data as they are
// then, after the caller's pre-snap.
ez::Stepper* session() {
  Lib& l = L();
  if (session_current()) return l.st;
  session_end();
  ensure_device();
  const long long n = nvox();
  if (!input_data || !decp_data) fail("input_data / decp_data not allocated");
  l.W = width;
  l.H = height;
  l.D = depth;
  l.planar = planar_mode();
  l.xi = bound;
  l.orig_host = input_data;
  hipStream_t q = ez::ctx_stream(l.work);
  h2d(l.orig.ensure(n), input_data, n, q);
  h2d(l.decp.ensure(n), decp_data, n, q);
  l.cnt.ensure(n);
  sync(q);
  l.shadow.assign(decp_data, decp_data + n);
  l.st = ez::stepper_new(l.work, l.planar, 0);
  ez::stepper_begin(l.st, l.orig.p, l.decp.p, l.cnt.p, width, height, depth, bound);
  return l.st;
}

// decp_data edited by the caller since the last synchronisation: the device
// copy is replaced and the correction's incremental state dropped.
void sync_in() {
  Lib& l = L();
  const long long n = nvox();
  if (std::memcmp(l.shadow.data(), decp_data, sizeof(double) * n) == 0) return;
  hipStream_t q = ez::ctx_stream(l.work);
  h2d(l.decp.p, decp_data, n, q);
  sync(q);
  l.shadow.assign(decp_data, decp_data + n);
  ez::stepper_decp_replaced(l.st);
}

// decp_data and delta_counter after a step that edits decp.
void sync_out() {
  Lib& l = L();
  const long long n = nvox();
  hipStream_t q = ez::ctx_stream(l.work);
  l.hcnt.resize(n);
  d2h(decp_data, l.decp.p, n, q);
  d2h(l.hcnt.data(), l.cnt.p, n, q);
  sync(q);
  std::memcpy(l.shadow.data(), decp_data, sizeof(double) * n);
  if ((long long)delta_counter.size() < n) delta_counter.resize(n, 0);
  for (long long i = 0; i < n; i++) delta_counter[i] = l.hcnt[i];
}

void counters() {
  int c[6];
  ez::stepper_counters(L().st, c);
  wrong_max_counter = c[0];
  wrong_max_counter_2 = c[1];
  wrong_min_counter = c[2];
  wrong_min_counter_2 = c[3];
  wrong_saddle_counter = c[4];
  wrong_saddle_counter_join = c[5];
}

// Every API call but the per-saddle ones starts a new generation.
struct Call {
  std::lock_guard<std::recursive_mutex> lk;
  Call() : lk(L().mu) { L().gen.fetch_add(1); }
};

// Runs an API body; errors are printed (the reference reports to std::cerr
// and returns), a missing GPU propagates.
template <class F>
auto guard(const char* name, F&& fn) -> decltype(fn()) {
  try {
    return fn();
  } catch (const std::bad_alloc&) {
    std::cerr << name << ": device allocation failed" << std::endl;
    throw;
  } catch (const ez::HipError& e) {
    std::cerr << name << ": " << e.what() << std::endl;
    throw std::runtime_error(std::string(name) + ": " + e.what());
  } catch (const std::range_error& e) {
    std::cerr << name << ": iteration cap reached in " << e.what() << std::endl;
    throw std::runtime_error(std::string(name) + ": iteration cap");
  }
}

// Labels of a caller field for the per-saddle label look-ups, once per
// generation.
const Lib::LabelCache& labels_of(const double* offset) {
  Lib& l = L();
  const unsigned long long g = l.gen.load();
  if (l.lab.gen.load(std::memory_order_acquire) == g && l.lab.field == offset) return l.lab;
  std::lock_guard<std::recursive_mutex> lk(l.mu);
  if (l.lab.gen.load() == g && l.lab.field == offset) return l.lab;
  ensure_device();
  const long long n = nvox();
  const double* d = upload_field(offset);
  int* ds = l.ia.ensure(n);
  int* as = l.ib.ensure(n);
  ez::field_labels(l.aux, planar_mode(), d, width, height, depth, ds, as);
  l.lab.ds.resize(n);
  l.lab.as.resize(n);
  hipStream_t q = aux_stream();
  d2h(l.lab.ds.data(), ds, n, q);
  d2h(l.lab.as.data(), as, n, q);
  sync(q);
  l.lab.field = offset;
  l.lab.gen.store(g, std::memory_order_release);
  return l.lab;
}

// The wrong-neighbour candidates of the correction's last CP pass, once per
// generation.
const Lib::CandCache& candidates() {
  Lib& l = L();
  const unsigned long long g = l.gen.load();
  if (l.cand.gen.load(std::memory_order_acquire) == g) return l.cand;
  std::lock_guard<std::recursive_mutex> lk(l.mu);
  if (l.cand.gen.load() == g) return l.cand;
  ez::Stepper* st = session();
  std::vector<int2> c;
  ez::stepper_candidates(st, c);
  l.cand.split.clear();
  l.cand.join.clear();
  for (const int2& e : c) (e.x & 1 ? l.cand.join : l.cand.split)[e.x >> 1] = (unsigned)e.y;
  l.cand.gen.store(g, std::memory_order_release);
  return l.cand;
}

int offset_of(int d) {
  return directions[3 * d] + width * (directions[3 * d + 1] + height * directions[3 * d + 2]);
}

bool in_grid(int v, int d) {
  const int x = v % width, y = (v / width) % height, z = v / (width * height);
  const int nx = x + directions[3 * d], ny = y + directions[3 * d + 1],
            nz = z + directions[3 * d + 2];
  return nx >= 0 && nx < width && ny >= 0 && ny < height && nz >= 0 && nz < depth;
}

bool larger_at(const double* f, int a, int b) {  // islarger, src/ExtreM.cpp:581-583
  return f[a] > f[b] || (f[a] == f[b] && a > b);
}

// compute_Max_for_Saddle / compute_Min_for_Saddle (src/ExtreM.cpp:2373-2403):
// the root label of each upper (lower) neighbour, in adjacency order, into
// the saddle's four slots of the map (the correction compares all of them
// on the device; the map holds the first four).
void saddle_slots(int saddle, const double* offset, bool up, int* map) {
  if (!map) return;
  const Lib::LabelCache& c = labels_of(offset);
  const int M = engine_dirs();
  int k = 0;
  for (int d = 0; d < M && k < 4; d++) {
    if (!in_grid(saddle, d)) continue;
    const int r = saddle + offset_of(d);
    if (up ? larger_at(offset, r, saddle) : larger_at(offset, saddle, r))
      map[saddle * 4LL + k++] = up ? c.ds[r] : c.as[r];
  }
}

// get_wrong_split_neighbors / get_wrong_join_neighbors (src/ExtreM.cpp:
// 2405-2450): the first candidate neighbour not yet flagged is flagged and
// appended.
void wrong_neighbor(int saddle, int& count, bool split) {
  const Lib::CandCache& c = candidates();
  const auto& m = split ? c.split : c.join;
  auto it = m.find(saddle);
  if (it == m.end()) return;
  int* flags = split ? wrong_neighbors : wrong_neighbors_ds;
  int* list = split ? wrong_neighbors_index : wrong_neighbors_ds_index;
  if (!flags || !list) fail("wrong_neighbors arrays not allocated");
  unsigned mask = it->second;
  while (mask) {
    const int d = __builtin_ctz(mask);
    mask &= mask - 1;
    const int r = saddle + offset_of(d);
    if (flags[r] == 0) {
      flags[r] = 1;
      list[count++] = r;
      return;
    }
  }
}

// vertex order positions of ids in field d_f (value, then id)
std::vector<int> positions_of(const double* d_f, const std::vector<int>& ids) {
  Lib& l = L();
  const long long n = nvox();
  hipStream_t q = aux_stream();
  int* sorted = l.ic.ensure(n);
  sort_by_value(d_f, nullptr, n, sorted);
  int* pos = l.id.ensure(n);
  k_rank<<<grid_for(n), 256, 0, q>>>(sorted, n, pos);
  std::vector<int> out(ids.size());
  if (!ids.empty()) {
    int* di = l.ie.ensure(ids.size());
    int* dout = l.ia.ensure(std::max<long long>(n, (long long)ids.size()));
    h2d(di, ids.data(), (long long)ids.size(), q);
    k_gather<<<grid_for(ids.size()), 256, 0, q>>>(pos, n, di, (long long)ids.size(), dout);
    d2h(out.data(), dout, (long long)ids.size(), q);
  }
  sync(q);
  return out;
}

}  // namespace ezc

using namespace ezc;

// ------------------------------------------------------------- the API ----
void Branch::print() const {
  std::cout << "Branch:" << std::endl << "  Vertices:" << std::endl;
  for (const auto& v : vertices)
    std::cout << "    Order: " << v.first << ", GlobalId: " << v.second << std::endl;
  if (parentBranch && !parentBranch->vertices.empty())
    std::cout << "  Parent Branch's first vertex GlobalId: " << parentBranch->vertices[0].second
              << std::endl;
  else
    std::cout << "  No Parent Branch" << std::endl;
}

extern "C" {

// saveArrayToBin, src/ExtreM.cpp:2918-2928: raw float64, no header.
void saveArrayToBin(const double* arr, size_t size, const std::string& filename) {
  std::ofstream out(filename, std::ios::binary);
  if (!out) {
    std::cerr << "Error opening file for writing: " << filename << std::endl;
    return;
  }
  out.write(reinterpret_cast<const char*>(arr), size * sizeof(double));
}

// getdata, src/ExtreM.cpp:94-140: reads the float64 field into input_data and
// sets bound = (max - min) * er.  The lossy field comes from the file named
// by EXTREMZ_DECP (raw float64 of the same size) when set, else from the
// uniform quantiser 2*bound*rint(orig / (2*bound)); the reference runs SZ3
// here, which this build does not include.  decp_data_copy receives a copy.
void getdata(const std::string& filename, const double er, int data_size, std::string type) {
  guard("getdata", [&] {
    Call call;
    session_end();
    if (!input_data || !decp_data) fail("input_data / decp_data not allocated");
    std::ifstream file(filename, std::ios::binary | std::ios::ate);
    if (!file) {
      std::cerr << "Cannot open file: " << filename << std::endl;
      return;
    }
    const std::streamsize size = file.tellg();
    file.seekg(0, std::ios::beg);
    if (size != static_cast<std::streamsize>(data_size * sizeof(double))) {
      std::cerr << "File size does not match expected data size." << std::endl;
      return;
    }
    file.read(reinterpret_cast<char*>(input_data), size);
    double lo = input_data[0], hi = input_data[0];
    for (int i = 1; i < data_size; i++) {
      lo = std::min(lo, input_data[i]);
      hi = std::max(hi, input_data[i]);
    }
    bound = (hi - lo) * er;
    if (const char* path = std::getenv("EXTREMZ_DECP"); path && *path) {
      std::ifstream d(path, std::ios::binary | std::ios::ate);
      if (!d || d.tellg() != size) fail(std::string("EXTREMZ_DECP file missing or of another size: ") + path);
      d.seekg(0, std::ios::beg);
      d.read(reinterpret_cast<char*>(decp_data), size);
    } else {
      ensure_device();
      Lib& l = L();
      hipStream_t q = aux_stream();
      double* dorig = l.f.ensure(data_size);
      double* dd = (double*)l.ka.ensure(data_size);
      h2d(dorig, input_data, data_size, q);
      k_quantise<<<grid_for(data_size), 256, 0, q>>>(dorig, dd, data_size, bound);
      d2h(decp_data, dd, data_size, q);
      sync(q);
    }
    if (decp_data_copy) std::copy(decp_data, decp_data + data_size, decp_data_copy);
    (void)type;
  });
}

// getSortedIndexes, src/ExtreM.cpp:1826-1838: vertex ids by ascending
// (value, id).
void getSortedIndexes(const double* vec, int* indexes) {
  guard("getSortedIndexes", [&] {
    Call call;
    ensure_device();
    const long long n = nvox();
    const double* d = upload_field(vec);
    int* out = L().ia.ensure(n);
    sort_by_value(d, nullptr, n, out);
    d2h(indexes, out, n, aux_stream());
    sync(aux_stream());
  });
}

// computeAdjacency, src/ExtreM.cpp:543-575.
void computeAdjacency() {
  guard("computeAdjacency", [&] {
    Call call;
    ensure_device();
    if (!adjacency) fail("adjacency not allocated");
    if (maxNeighbors < 1 || maxNeighbors > 14) fail("maxNeighbors must be in [1, 14]");
    const long long n = nvox();
    hipStream_t q = aux_stream();
    int* d = L().ia.ensure(n * maxNeighbors);
    k_adjacency<<<grid_for(n), 256, 0, q>>>(grid(), maxNeighbors, d);
    d2h(adjacency, d, n * maxNeighbors, q);
    sync(q);
  });
}

// get_vertex_traingle, src/ExtreM.cpp:1585-1595 with getVerticesFromTriangleID1
// (:597-627): the triangles of the z = 0 plane listed per vertex.
void get_vertex_traingle() {
  Call call;
  const long long faces = 2LL * (width - 1) * (height - 1);
  if ((long long)vertex_cells.size() < (long long)width * height)
    vertex_cells.resize((size_t)width * height);
  for (long long i = 0; i < faces; i++) {
    const long long base = i / 2;
    const int y = (int)((base % ((long long)(width - 1) * (height - 1))) / (width - 1));
    const int x = (int)(base % (width - 1));
    int v[3];
    if (i % 2 == 0) {
      v[0] = x + y * width;
      v[1] = x + 1 + y * width;
      v[2] = x + (y + 1) * width;
    } else {
      v[0] = x + 1 + (y + 1) * width;
      v[1] = x + (y + 1) * width;
      v[2] = x + 1 + y * width;
    }
    for (int k = 0; k < 3; k++) vertex_cells[v[k]].push_back((int)i);
  }
}

// classifyAllVertices, src/ExtreM.cpp:1088-1097 (classifyVertex1 :703-891):
// the types by the link LUT of the correction's adjacency, the star lists
// over the first maxNeighbors directions.
void classifyAllVertices(int* vt, const double* heightMap, int* lower, int* upper, int type) {
  guard("classifyAllVertices", [&] {
    Call call;
    ensure_device();
    const long long n = nvox();
    Lib& l = L();
    hipStream_t q = aux_stream();
    const double* d = upload_field(heightMap);
    uint8_t* t = l.u8a.ensure(n);
    uint16_t* lm = l.u16.ensure(n);
    sync(q);
    ez::field_classify(l.aux, planar_mode(), d, width, height, depth, t, lm);
    if (vt) {
      int* w = l.ia.ensure(n);
      k_widen<<<grid_for(n), 256, 0, q>>>(t, w, n);
      d2h(vt, w, n, q);
    }
    if (lower || upper) {
      const long long m = n * (maxNeighbors + 1);
      int* dl = lower ? l.ib.ensure(m) : nullptr;
      int* du = upper ? l.ic.ensure(m) : nullptr;
      k_stars<<<grid_for(n), 256, 0, q>>>(grid(), maxNeighbors, d, dl, du);
      if (lower) d2h(lower, dl, m, q);
      if (upper) d2h(upper, du, m, q);
    }
    sync(q);
    (void)type;
  });
}

// ComputeDescendingManifold / ComputeAscendingManifold (src/ExtreM.cpp:
// 144-183, 502-541): fully compressed root labels.  On decp_data during a
// correction these are the correction's own (its CP pass of decp runs when
// decp changed since the last one); on any other field a label pass.
static int manifold(const double* offset, int* out, bool descending) {
  return guard(descending ? "ComputeDescendingManifold" : "ComputeAscendingManifold", [&] {
    Call call;
    ensure_device();
    const long long n = nvox();
    Lib& l = L();
    if (offset == decp_data && session_current()) {
      sync_in();
      ez::stepper_cp(l.st);
      const int *ds, *as;
      ez::stepper_labels(l.st, &ds, &as);
      hipStream_t q = ez::ctx_stream(l.work);
      d2h(out, descending ? ds : as, n, q);
      sync(q);
      return 0;
    }
    const double* d = upload_field(offset);
    int* ds = l.ia.ensure(n);
    int* as = l.ib.ensure(n);
    ez::field_labels(l.aux, planar_mode(), d, width, height, depth, ds, as);
    d2h(out, descending ? ds : as, n, aux_stream());
    sync(aux_stream());
    return 0;
  });
}

int ComputeDescendingManifold(const double* offset, const int* vt, int* DS_M) {
  (void)vt;
  return manifold(offset, DS_M, true);
}

int ComputeAscendingManifold(const double* offset, const int* vt, int* AS_M) {
  (void)vt;
  return manifold(offset, AS_M, false);
}

// mappath, src/ExtreM.cpp:402-437: both label arrays.
void mappath(int* DS_M, int* AS_M, const double* offset) {
  guard("mappath", [&] {
    Call call;
    ensure_device();
    const long long n = nvox();
    Lib& l = L();
    const double* d = upload_field(offset);
    int* ds = l.ia.ensure(n);
    int* as = l.ib.ensure(n);
    ez::field_labels(l.aux, planar_mode(), d, width, height, depth, ds, as);
    d2h(DS_M, ds, n, aux_stream());
    d2h(AS_M, as, n, aux_stream());
    sync(aux_stream());
  });
}

// computePathCompressionSingle, src/ExtreM.cpp:439-495: one label array.
int computePathCompressionSingle(int* const segmentation, const bool computeAscending,
                                 const double* const offset) {
  return guard("computePathCompressionSingle", [&] {
    Call call;
    ensure_device();
    const long long n = nvox();
    Lib& l = L();
    const double* d = upload_field(offset);
    int* ds = l.ia.ensure(n);
    int* as = l.ib.ensure(n);
    ez::field_labels(l.aux, planar_mode(), d, width, height, depth, ds, as);
    d2h(segmentation, computeAscending ? as : ds, n, aux_stream());
    sync(aux_stream());
    return 0;
  });
}

// computeCP, src/ExtreM.cpp:1920-2040 (findAscPaths / findDesPaths
// :1597-1679): the CP lists of vertex_type_tmp sorted as the reference
// sorts them (2-saddles and minima by descending, 1-saddles and maxima by
// ascending (value, id)), their triplets of ranks over the correction's
// adjacency, and with translation the argmin of offset.
void computeCP(std::vector<std::pair<int, int>>& maximaTriplets,
               std::vector<std::pair<int, int>>& minimumTriplets,
               std::vector<std::array<int, 46>>& s2T, std::vector<std::array<int, 46>>& s1T,
               const double* offset, const int* desManifold, const int* ascManifold,
               const int* vertex_type_tmp, std::vector<int>& s2, std::vector<int>& s1,
               std::vector<int>& mn, std::vector<int>& mx, int& gmin, int translation) {
  guard("computeCP", [&] {
    Call call;
    ensure_device();
    (void)maximaTriplets;
    (void)minimumTriplets;
    const long long n = nvox();
    Lib& l = L();
    hipStream_t q = aux_stream();
    const Grid g = grid();
    const double* d = upload_field(offset);
    int* dt = l.ia.ensure(n);
    int* dsm = l.ib.ensure(n);
    int* asm_ = l.ic.ensure(n);
    h2d(dt, vertex_type_tmp, n, q);
    h2d(dsm, desManifold, n, q);
    h2d(asm_, ascManifold, n, q);
    uint8_t *f2 = l.u8a.ensure(n), *f1 = l.u8b.ensure(n), *fx = l.u8c.ensure(n),
            *fn = l.u8d.ensure(n);
    k_cp_flags<<<grid_for(n), 256, 0, q>>>(dt, n, f2, f1, fx, fn);
    int* lists = l.lists.ensure(4 * n + 4);
    int* tmp = l.id.ensure(n);
    int* rank = l.rank.ensure(n);
    long long cnt[4];
    int* out[4] = {lists, lists + n, lists + 2 * n, lists + 3 * n};  // S2, S1, MX, MN
    const uint8_t* fl[4] = {f2, f1, fx, fn};
    const bool desc[4] = {true, false, false, true};
    std::vector<int>* hv[4] = {&s2, &s1, &mx, &mn};
    for (int k = 0; k < 4; k++) {
      cnt[k] = select_flagged(fl[k], n, tmp);
      if (!cnt[k]) continue;
      if (desc[k]) {
        sort_by_value(d, tmp, cnt[k], l.ie.ensure(n));
        k_reverse<<<grid_for(cnt[k]), 256, 0, q>>>(l.ie.p, out[k], cnt[k]);
      } else {
        sort_by_value(d, tmp, cnt[k], out[k]);
      }
    }
    for (int k = 0; k < 4; k++) {
      hv[k]->resize(cnt[k]);
      d2h(hv[k]->data(), out[k], cnt[k], q);
    }
    // 2-saddles: ranks of the maxima their upper neighbours ascend to;
    // 1-saddles: ranks of the minima their lower neighbours descend to
    const int M = engine_dirs();
    s2T.clear();
    s2T.resize(cnt[0]);
    s1T.clear();
    s1T.resize(cnt[1]);
    sync(q);
    int* trip = l.trip.ensure(46 * std::max(cnt[0], cnt[1]) + 1);
    for (int side = 0; side < 2; side++) {
      const long long ns = cnt[side];
      if (!ns) continue;
      const int* ext = side == 0 ? out[2] : out[3];
      const long long ne = side == 0 ? cnt[2] : cnt[3];
      k_fill<<<grid_for(n), 256, 0, q>>>(rank, -1, n);
      if (ne) k_rank<<<grid_for(ne), 256, 0, q>>>(ext, ne, rank);
      k_triplets<<<grid_for(ns), 256, 0, q>>>(g, M, out[side], ns, d, side == 0 ? dsm : asm_,
                                              rank, side == 0, trip);
      auto& T = side == 0 ? s2T : s1T;
      d2h(&T[0][0], trip, 46 * ns, q);
      sync(q);
    }
    sync(q);
    if (translation != 0) {  // getSortedIndexes(offset)[0]
      int* order = l.ie.ensure(n);
      sort_by_value(d, nullptr, n, order);
      d2h(&gmin, order, 1, q);
      sync(q);
    }
  });
}

// findAscPaths, src/ExtreM.cpp:1597-1636: the triplets of the listed
// 2-saddles (ranks through ascManifold, here the rank array of the maxima).
void findAscPaths(std::vector<std::array<int, 46>>& T, std::vector<int>& maximaLocalToGlobal,
                  std::vector<int>& saddlesLocalToGlobal, const int* s2, const int* mx,
                  const double* offset, const int* desManifold, const int* ascManifold,
                  int& counter_tmp) {
  guard("findAscPaths", [&] {
    Call call;
    ensure_device();
    (void)maximaLocalToGlobal;
    (void)mx;
    (void)counter_tmp;
    const long long n = nvox();
    const long long ns = (long long)saddlesLocalToGlobal.size();
    if ((long long)T.size() < ns) T.resize(ns);
    if (!ns) return;
    Lib& l = L();
    hipStream_t q = aux_stream();
    const double* d = upload_field(offset);
    int* dsm = l.ib.ensure(n);
    int* rank = l.rank.ensure(n);
    int* list = l.ic.ensure(ns);
    h2d(dsm, desManifold, n, q);
    h2d(rank, ascManifold, n, q);
    h2d(list, s2, ns, q);
    int* trip = l.trip.ensure(46 * ns);
    k_triplets<<<grid_for(ns), 256, 0, q>>>(grid(), engine_dirs(), list, ns, d, dsm, rank, 1,
                                            trip);
    std::vector<int> h(46 * ns);
    d2h(h.data(), trip, 46 * ns, q);
    sync(q);
    for (long long i = 0; i < ns; i++) {  // [45] is computeCP's, not findAscPaths'
      const int keep = T[i][45];
      std::copy(h.begin() + 46 * i, h.begin() + 46 * i + 46, T[i].begin());
      T[i][45] = keep;
    }
  });
}

// sortAndRemoveDuplicates, src/ExtreM.cpp:585-595.
void sortAndRemoveDuplicates(std::array<int, 46>& triplet) {
  const int c = std::max(0, std::min(triplet[44], 44));
  std::sort(triplet.begin(), triplet.begin() + c);
  int u = 1;
  for (int p = 1; p < c; p++)
    if (triplet[p - 1] != triplet[p]) triplet[u++] = triplet[p];
  triplet[44] = u;
}

// computelargestSaddlesForMax / computesmallestSaddlesForMin
// (src/ExtreM.cpp:2309-2371): per extremum (vertex id key), the first
// triplet (list position) that reaches it; the list size when none does or
// for the last (global) extremum.
static void first_saddles(int nExt, const std::vector<std::array<int, 46>>& T,
                          std::unordered_map<int, int>& out, const std::vector<int>* ext) {
  out.clear();
  if (!ext) return;
  const int n = (int)std::min<size_t>(nExt, ext->size());
  std::vector<int> best(n, (int)T.size());
  const int global = nExt - 1;
  for (int i = 0; i < (int)T.size(); i++) {
    const auto& t = T[i];
    for (int p = 0; p < t[44] && p < 44; p++) {
      const int e = t[p];
      if (e == global || e < 0 || e >= n) continue;
      best[e] = std::min(best[e], i);
    }
  }
  out.reserve(n);
  for (int i = 0; i < n; i++) out[(*ext)[i]] = best[i];
}

void computelargestSaddlesForMax(const int nMax, const std::vector<std::array<int, 46>>& T,
                                 std::unordered_map<int, int>& out, const std::vector<int>* mx) {
  Call call;
  first_saddles(nMax, T, out, mx);
}

void computesmallestSaddlesForMin(const int nMin, const std::vector<std::array<int, 46>>& T,
                                  std::unordered_map<int, int>& out, const std::vector<int>* mn) {
  Call call;
  first_saddles(nMin, T, out, mn);
}

// c_loops, src/ExtreM.cpp:1505-1583: until no false critical point is left.
// Afterwards dec_vertex_type holds decp's types (the original ones) and the
// decp star lists are recomputed.
void c_loops(std::string type) {
  guard("c_loops", [&] {
    Call call;
    check_split(type);
    ez::Stepper* st = session();
    sync_in();
    ez::stepper_c_loops(st);
    sync_out();
    count_f_max = 0;
    count_f_min = 0;
    count_f_saddle = 0;
    Lib& l = L();
    const long long n = nvox();
    hipStream_t q = ez::ctx_stream(l.work);
    if (dec_vertex_type) {
      int* w = l.ie.ensure(n);
      k_widen<<<grid_for(n), 256, 0, q>>>(ez::stepper_types(st), w, n);
      d2h(dec_vertex_type, w, n, q);
    }
    if (dec_lowerStars || dec_upperStars) {
      const long long m = n * (maxNeighbors + 1);
      int* dl = dec_lowerStars ? l.ib.ensure(m) : nullptr;
      int* du = dec_upperStars ? l.ic.ensure(m) : nullptr;
      k_stars<<<grid_for(n), 256, 0, q>>>(grid(), maxNeighbors, l.decp.p, dl, du);
      if (dl) d2h(dec_lowerStars, dl, m, q);
      if (du) d2h(dec_upperStars, du, m, q);
    }
    sync(q);
  });
}

// r_loops, src/ExtreM.cpp:2577-2593: fixpath over the wrong neighbours the
// caller collected, then apply_delta.
void r_loops(std::string type) {
  guard("r_loops", [&] {
    Call call;
    check_split(type);
    ez::Stepper* st = session();
    sync_in();
    const int n2 = wrong_neighbors_index ? number_of_false_cases : 0;
    const int n1 = wrong_neighbors_ds_index ? number_of_false_cases1 : 0;
    ez::stepper_r_loops(st, wrong_neighbors_index, n2, wrong_neighbors_ds_index, n1);
    sync_out();
  });
}

// fixpath, src/ExtreM.cpp:2452-2575: one wrong neighbour's proposal into
// d_deltaBuffer.  Every exported step re-initialises the delta buffer
// before it applies (init_delta, :1099-1105), so a proposal made outside
// r_loops never reaches decp: the call returns the reference's 0 and leaves
// decp unchanged.  r_loops runs the path walks on the device.
int fixpath(int i, int direction, std::string type) {
  Call call;
  (void)i;
  (void)direction;
  (void)type;
  return 0;
}

static void s_loops_side(const char* name, const std::string& type, int side, int fix) {
  guard(name, [&] {
    Call call;
    if (fix) check_split(type);
    ez::Stepper* st = session();
    sync_in();
    if (fix == 2) ez::stepper_saddle_loops(st, side, 1);
    else if (fix == 3) ez::stepper_saddle_loops(st, side, 0);
    else ez::stepper_s_loops(st, side, fix);
    const int keep[6] = {wrong_max_counter, wrong_max_counter_2, wrong_min_counter,
                         wrong_min_counter_2, wrong_saddle_counter, wrong_saddle_counter_join};
    counters();
    // only the counters the reference function sets change
    const bool wi = fix <= 1;
    if (!(wi && side == 2)) {
      wrong_max_counter = keep[0];
      wrong_max_counter_2 = keep[1];
    }
    if (!(wi && side == 1)) {
      wrong_min_counter = keep[2];
      wrong_min_counter_2 = keep[3];
    }
    if (!(!wi && side == 2)) wrong_saddle_counter = keep[4];
    if (!(!wi && side == 1)) wrong_saddle_counter_join = keep[5];
    if (fix == 1 || fix == 2) sync_out();
  });
}

// s_loops / s_loops_join (src/ExtreM.cpp:2838-2886).
void s_loops(std::string type) { s_loops_side("s_loops", type, 2, 1); }
void s_loops_join(std::string type) { s_loops_side("s_loops_join", type, 1, 1); }
// saddle_loops / saddle_loops_join (src/ExtreM.cpp:2888-2916).
void saddle_loops(std::string type) { s_loops_side("saddle_loops", type, 2, 2); }
void saddle_loops_join(std::string type) { s_loops_side("saddle_loops_join", type, 1, 2); }
// get_wrong_index_max / _min / _saddles / _saddles_join (src/ExtreM.cpp:
// 2595-2715): the counters of the correction's last CP state.
void get_wrong_index_max() { s_loops_side("get_wrong_index_max", "split", 2, 0); }
void get_wrong_index_min() { s_loops_side("get_wrong_index_min", "split", 1, 0); }
void get_wrong_index_saddles() { s_loops_side("get_wrong_index_saddles", "split", 2, 3); }
void get_wrong_index_saddles_join() {
  s_loops_side("get_wrong_index_saddles_join", "split", 1, 3);
}

void compute_Max_for_Saddle(int saddle, const double* offset) {
  guard("compute_Max_for_Saddle", [&] { saddle_slots(saddle, offset, true, or_saddle_max_map); });
}

void compute_Min_for_Saddle(int saddle, const double* offset) {
  guard("compute_Min_for_Saddle", [&] { saddle_slots(saddle, offset, false, or_saddle_min_map); });
}

void get_wrong_split_neighbors(int saddle, int& num_false_cases, const double* offset) {
  (void)offset;
  guard("get_wrong_split_neighbors", [&] { wrong_neighbor(saddle, num_false_cases, true); });
}

void get_wrong_join_neighbors(int saddle, int& num_false_cases, const double* offset) {
  (void)offset;
  guard("get_wrong_join_neighbors", [&] { wrong_neighbor(saddle, num_false_cases, false); });
}

// get_false_criticle_points, src/ExtreM.cpp:1324-1383, over the type and
// star arrays: counts and lists (in vertex order).
void get_false_criticle_points() {
  guard("get_false_criticle_points", [&] {
    Call call;
    ensure_device();
    if (!vertex_type || !dec_vertex_type) fail("vertex_type / dec_vertex_type not allocated");
    const long long n = nvox();
    Lib& l = L();
    hipStream_t q = aux_stream();
    const long long m = n * (maxNeighbors + 1);
    const bool stars = lowerStars && upperStars && dec_lowerStars && dec_upperStars;
    Dev<int> sl, su, sdl, sdu;  // star arrays (freed below)
    int *t = l.ia.ensure(n), *dt = l.ib.ensure(n);
    h2d(t, vertex_type, n, q);
    h2d(dt, dec_vertex_type, n, q);
    if (stars) {
      h2d(sl.ensure(m), lowerStars, m, q);
      h2d(su.ensure(m), upperStars, m, q);
      h2d(sdl.ensure(m), dec_lowerStars, m, q);
      h2d(sdu.ensure(m), dec_upperStars, m, q);
    }
    int *cx = l.ic.ensure(n + 1), *cn = l.id.ensure(n + 1), *cs = l.ie.ensure(n + 1);
    k_false_counts<<<grid_for(n), 256, 0, q>>>(n, maxNeighbors, t, dt, sl.p, su.p, sdl.p, sdu.p,
                                              cx, cn, cs);
    int* off = l.rank.ensure(n + 1);
    int* dst[3] = {all_max, all_min, all_saddle};
    int* cnts[3] = {cx, cn, cs};
    int total[3];
    for (int k = 0; k < 3; k++) {
      size_t bytes = 0;
      k_fill<<<1, 64, 0, q>>>(cnts[k] + n, 0, 1);
      hip_check(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnts[k], off, (int)(n + 1), q),
                "scan size");
      uint8_t* tmp = l.tmp.ensure(bytes);
      hip_check(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, cnts[k], off, (int)(n + 1), q),
                "scan");
      d2h(&total[k], off + n, 1, q);
      sync(q);
      if (dst[k] && total[k]) {
        int* o = l.lists.ensure(std::max<long long>(total[k], 4 * n + 4));
        k_false_write<<<grid_for(n), 256, 0, q>>>(n, cnts[k], off, o);
        d2h(dst[k], o, total[k], q);
        sync(q);
      }
    }
    for (Dev<int>* b : {&sl, &su, &sdl, &sdu})
      if (b->p) hip_check(hipFree(b->p), "hipFree");
    count_f_max = total[0];
    count_f_min = total[1];
    count_f_saddle = total[2];
  });
}

// classifyVertex, src/ExtreM.cpp:1727-1824, for one vertex over the
// caller's adjacency and labels: 0 when no neighbour is lower, 4 when none
// is higher, 5 with one distinct label on each side; otherwise types
// receives 2 and 1 (the reference's link-component count is always 5,
// computeNumberOfLinkComponents below) and it returns 5.
int classifyVertex(const int v, const double* offset, const int* desManifold,
                   const int* ascManifold, int* types) {
  Call call;
  types[0] = -1;
  types[1] = -1;
  // label runs: a label different from the last one stored starts a new entry
  int up_last = -1, lo_last = -1, nup = 0, nlo = 0, nu2 = 0, nl2 = 0;
  for (int k = 0; k < maxNeighbors; k++) {
    const int u = adjacency[(long long)maxNeighbors * v + k];
    if (u == -1) continue;
    if (larger_at(offset, u, v)) {
      nu2++;
      if (up_last != desManifold[u]) {
        up_last = desManifold[u];
        nup++;
      }
    } else if (larger_at(offset, v, u)) {
      nl2++;
      if (lo_last != ascManifold[u]) {
        lo_last = ascManifold[u];
        nlo++;
      }
    }
  }
  if (nlo == 0) {
    types[0] = 0;
    return 0;
  }
  if (nup == 0) {
    types[0] = 4;
    return 4;
  }
  if (nup == 1 && nlo == 1) {
    types[0] = 5;
    return 5;
  }
  if (computeNumberOfLinkComponents(nullptr, nu2) > 1) {
    if (types[0] == -1) types[0] = 2;
    else types[1] = 2;
  }
  if (computeNumberOfLinkComponents(nullptr, nl2) > 1) {
    if (types[0] == -1) types[0] = 1;
    else types[1] = 1;
  }
  return 5;
}

// computeNumberOfLinkComponents, src/ExtreM.cpp:1810-1824: the reference's
// body builds a map of the link vertices and returns 5 without counting.
int computeNumberOfLinkComponents(const int* linkVertices, const int nLinkVertices) {
  (void)linkVertices;
  (void)nLinkVertices;
  return 5;
}

// reorderVectorByTemparray, src/ExtreM.cpp:2954-2975: vec reordered by
// ascending temparray value.
void reorderVectorByTemparray(std::vector<std::pair<int, int>>& vec,
                              const std::vector<int>& temparray) {
  Call call;
  std::vector<size_t> idx(vec.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(),
            [&](size_t a, size_t b) { return temparray[a] < temparray[b]; });
  std::vector<std::pair<int, int>> out(vec.size());
  for (size_t i = 0; i < idx.size(); i++) out[i] = vec[idx[i]];
  vec = out;
}

// constructPersistencePairs, src/ExtreM.cpp:2132-2249, over the caller's
// rank triplets (edited in place as the reference does): rounds pairing each
// maximum that is the lowest of its largest saddle, pointer compression,
// relabelling.
int constructPersistencePairs(std::vector<std::pair<int, int>>& pairs,
                              std::vector<std::pair<int, int>>& maximaTriplets,
                              std::vector<std::array<int, 46>>& T, const int* Maximum,
                              const int* s2, const int nMax, int gmin) {
  Call call;
  if (nMax <= 0) return 1;
  if ((int)pairs.size() < nMax) pairs.resize(nMax);
  if ((int)maximaTriplets.size() < nMax) maximaTriplets.resize(nMax);
  std::vector<int> ptr(nMax);
  std::iota(ptr.begin(), ptr.end(), 0);
  const int gmax = nMax - 1;
  for (bool changed = true; changed;) {
    changed = false;
    std::vector<int> largest(nMax, (int)T.size());
    for (int i = 0; i < (int)T.size(); i++)
      for (int p = 0; p < T[i][44] && p < 44; p++) {
        const int m = T[i][p];
        if (m != gmax && m >= 0 && m < nMax) largest[m] = std::min(largest[m], i);
      }
    std::vector<int> active;
    for (int m = 0; m < nMax; m++) {
      const int s = largest[m];
      if (s < (int)T.size() && T[s][0] == m) {
        changed = true;
        pairs[m] = {s2[s], Maximum[m]};
        const int top = T[s][std::max(0, T[s][44] - 1)];
        ptr[m] = top;
        maximaTriplets[m] = {s, top};
        active.push_back(m);
      }
    }
    while (!active.empty()) {
      size_t k = 0;
      for (int m : active) {
        int& p = ptr[m];
        p = ptr[p];
        if (p != ptr[p]) active[k++] = m;
      }
      active.resize(k);
    }
    for (auto& t : T) {
      for (int r = 0; r < t[44] && r < 44; r++)
        if (t[r] >= 0 && t[r] < nMax) t[r] = ptr[t[r]];
      sortAndRemoveDuplicates(t);
      if (t[44] == 1) t[44] = 0;
    }
  }
  pairs[pairs.size() - 1] = {gmin, Maximum[nMax - 1]};
  maximaTriplets[maximaTriplets.size() - 1] = {gmin, (int)maximaTriplets.size() - 1};
  return 1;
}

// compute_MergeT, src/ExtreM.cpp:2930-2952: the split tree of offset with
// its own types, labels and value order (computed on the GPU), as the
// branches of constructMergeTree (:1863-1918) in ascending (value, id) order
// of the maxima: (order, vertex) pairs in descending order, the pair saddle
// and the saddles of child branches, (-1, root marker) for a root branch.
// maximaTriplets[b] = (index of b's saddle in saddles2, parent branch).
void compute_MergeT(std::vector<Branch>& branches, const std::vector<int> s2,
                    const std::vector<int> mx, std::vector<std::pair<int, int>>& maximaTriplets,
                    std::vector<std::array<int, 46>>& T, const double* offset, const int data_size,
                    const int gmin) {
  guard("compute_MergeT", [&] {
    Call call;
    ensure_device();
    (void)T;
    (void)data_size;
    Lib& l = L();
    const double* d = upload_field(offset);
    sync(aux_stream());
    ez::TreeOut t;
    ez::field_split_tree(l.aux, planar_mode(), d, width, height, depth, t);
    const int nmax = (int)t.max_v.size();
    if (nmax != (int)mx.size())
      std::cerr << "compute_MergeT: " << mx.size() << " maxima given, the field has " << nmax
                << std::endl;
    std::unordered_map<int, int> rank_of, sad_index;
    for (int r = 0; r < nmax; r++) rank_of[t.max_v[r]] = r;
    for (int i = 0; i < (int)s2.size(); i++) sad_index.emplace(s2[i], i);
    // order positions of every vertex a branch lists
    std::vector<int> ids;
    for (int b = 0; b < nmax; b++)
      for (int v : t.branch[b]) ids.push_back(v);
    const std::vector<int> pos = positions_of(d, ids);
    maximaTriplets.resize(nmax);
    branches.resize(nmax);
    size_t k = 0;
    for (int b = 0; b < nmax; b++) {
      const int parent = rank_of.count(t.parent_v[b]) ? rank_of[t.parent_v[b]] : b;
      const bool root = parent == b;
      int sad = 0;
      if (b == nmax - 1) sad = gmin;
      else if (sad_index.count(t.pair_v[b])) sad = sad_index[t.pair_v[b]];
      maximaTriplets[b] = {sad, parent};
      const auto& bv = t.branch[b];
      for (size_t j = 0; j < bv.size(); j++, k++) {
        const bool marker = root && j + 1 == bv.size();
        if (marker) branches[b].vertices.emplace_back(-1, b == nmax - 1 ? gmin : bv[j]);
        else branches[b].vertices.emplace_back(pos[k], bv[j]);
      }
      if (!root) branches[b].parentBranch = &branches[parent];
    }
  });
}

}  // extern "C"
