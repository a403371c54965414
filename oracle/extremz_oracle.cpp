// extremz_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the ExtreM correction path of the reference
// (control flow of main.cpp:65-446, rules of src/ExtreM.cpp), lifted to 3-D
// with the 14-neighbour Freudenthal stencil and the LUT classification of
// src/ExtreM_cuda.cu:3720-3853.  It is the checker for the HIP engine in
// extrem_z_amd/: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
// of bench.py may load it.  The product never links or calls it.
//
// Where the reference is undefined or order-dependent, the restatement makes
// one deterministic choice (tagged D1..D13 below and listed in DESIGN.md):
//   D1  14-neighbour stencil (src/ExtreM_cuda.cu:155-162) instead of
//       maxNeighbors=6 (main.cpp:20).
//   D2  LUT classification (src/ExtreM_cuda.cu:3835-3853) instead of the
//       z=0 triangle list (main.cpp:56, src/ExtreM.cpp:767 is UB for 3-D).
//   D3  14 label slots per saddle instead of 4 (main.cpp:142,155).
//   D4  CPU triplet semantics: dedupe, no persistence threshold.
//   D5  delta proposals are max-reduced (the sequential meaning of
//       c_loop's `d > oldValue` + swap, src/ExtreM.cpp:1134-1165).
//   D6  serial greedy orders (saddle id ascending, neighbours in adjacency
//       order) are kept literally here; the GPU reproduces them.
//   D7  manifold labels are fully compressed roots for every vertex
//       (src/ExtreM.cpp:166,524 gate compression on `vertex_type[3]`).
//   D8  tree_type is always "split" (main.cpp:63), so every apply snaps to
//       orig - xi, and the join saddle fix raises (src/ExtreM.cpp:2826).
//   D9  lower stars are compared as direction bitmasks (same order as the
//       reference's adjacency lists, so list equality == mask equality).
//   D10 triplets store the extremum's vertex id, ordered by (value, id); this
//       equals the reference's rank order for every extremum in the CP list.
//       A root outside the (possibly stale) CP list is ordered by the same key
//       (the reference reads uninitialised tempArray there) and is never a
//       key of the largest/smallest-saddle maps.
//   D11 an empty triplet gives NONE (-1) for its first/last entry.
//   D12 a largest/smallest-saddle index equal to the list size (no adjacent
//       saddle; always the case for the global extremum) reads as NONE
//       instead of one element past the end of the vector, and a pair with a
//       NONE side is no mismatch: it is neither counted nor fixed (the
//       reference reads past both vectors there; counting it would stall
//       the loop on a pair that no fix addresses, e.g. when decp's highest
//       maximum is not the original global maximum).
//   D13 fix_wrong_index_saddle is skipped when either index is NONE (kept as
//       a guard; D12 no longer produces such pairs).
//   D14 the split-tree check ranks value ties by vertex id
//       (getSortedPositions, src/ExtreM.cpp:1840-1860, leaves them to std::sort).
//   D15 the split tree of each field uses that field's own types, labels and
//       value order (main.cpp:459-461 sorts the original's lists by decp).
//   D16 a triplet end that is NONE (an empty triplet, D11) on either side of
//       a get_wrong_index_max/min comparison is no mismatch.  Only the
//       persistence filter empties a triplet; the GPU fork skips a saddle
//       whose original triplet is empty (src/ExtreM_cuda.cu:3115, 3146) and
//       reads stale entries for an empty decp one.
//   LUT extension: fields with an axis of length 1 give boundary classes the
//   reference has no LUT segment for; they use the same link-component rule.
//   Iteration caps turn the reference's unbounded loops into an error code.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kNbr = 14;
constexpr int NONE = -1;

// src/ExtreM_cuda.cu:155-162 (also src/ExtreM.cpp:16-21).
const int kDir[kNbr][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},
    {0, 0, -1}, {-1, 1, 0}, {1, -1, 0}, {0, 1, 1},  {0, -1, -1},
    {-1, 0, 1}, {1, 0, -1}, {-1, 1, 1}, {1, -1, -1}};

// Boundary keys in LUT.bin segment order, src/ExtreM_cuda.cu:84-91.
const int kRefKeys[27] = {277,  153,   1050,  2085,  102,   554,  413,
                          1179, 2357,  1594,  2151,  622,   5462, 10921,
                          5463, 5470,  5494,  10937, 10925, 10923, 5599,
                          7543, 11197, 6014,  11963, 10991, 16383};

// ---------------------------------------------------------------- LUT ----
// Link edge between neighbours a and b: both in the stencil and b - a is a
// stencil direction (generateLookupTable, src/ExtreM_cuda.cu:3280-3318).
bool is_dir(int dx, int dy, int dz) {
  for (int i = 0; i < kNbr; i++)
    if (kDir[i][0] == dx && kDir[i][1] == dy && kDir[i][2] == dz) return true;
  return false;
}

int uf_find(int* p, int x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}

// Type for one boundary key and one upper-bit config (bit k = k-th existing
// neighbour is upper), calculateLUT src/ExtreM_cuda.cu:263-467.
int link_type(int key, int config) {
  int ex[kNbr], m = 0;
  for (int d = 0; d < kNbr; d++)
    if (key >> d & 1) ex[m++] = d;
  int par[kNbr];
  for (int k = 0; k < m; k++) par[k] = k;
  for (int a = 0; a < m; a++)
    for (int b = a + 1; b < m; b++) {
      int da = ex[a], db = ex[b];
      if (!is_dir(kDir[db][0] - kDir[da][0], kDir[db][1] - kDir[da][1],
                  kDir[db][2] - kDir[da][2]))
        continue;
      if (((config >> a) & 1) != ((config >> b) & 1)) continue;
      int ra = uf_find(par, a), rb = uf_find(par, b);
      if (ra != rb) par[ra] = rb;
    }
  int lower = 0, upper = 0;
  for (int k = 0; k < m; k++)
    if (uf_find(par, k) == k) ((config >> k) & 1 ? upper : lower)++;
  // src/ExtreM_cuda.cu:458-464 (same rule as src/ExtreM.cpp:880-886)
  if (lower == 0 && upper == 1) return 0;
  if (lower == 1 && upper == 0) return 4;
  if (lower == 1 && upper == 1) return 5;
  if (lower > 1 && upper > 1) return 3;
  if (lower > 1) return 1;
  if (upper > 1) return 2;
  return 5;
}

struct Lut {
  // key (14-bit neighbour-existence mask) -> segment offset, -1 if absent
  std::vector<int> offset = std::vector<int>(1 << kNbr, -1);
  std::vector<int> table;
  int ref_size = 0;  // entries covered by the 27 reference keys
  Lut() {
    for (int key : kRefKeys) add(key);
    ref_size = (int)table.size();
    // extension classes (axes of length 1 / 2 combos not in the reference)
    for (int ax = 0; ax < 4; ax++)
      for (int ay = 0; ay < 4; ay++)
        for (int az = 0; az < 4; az++) {
          int st[3] = {ax, ay, az}, key = 0;
          for (int d = 0; d < kNbr; d++) {
            bool ok = true;
            for (int c = 0; c < 3; c++) {
              int s = kDir[d][c];
              if (s < 0 && !(st[c] == 0 || st[c] == 2)) ok = false;
              if (s > 0 && !(st[c] == 0 || st[c] == 1)) ok = false;
            }
            if (ok) key |= 1 << d;
          }
          if (offset[key] < 0) add(key);
        }
    // the CPU path's plane adjacency: +-x / +-y of each depth-1 class; the
    // triangle links of classifyVertex1 (src/ExtreM.cpp:760-818) join +x
    // with +y and -x with -y, which is link_type's rule on these keys
    for (int ax = 0; ax < 4; ax++)
      for (int ay = 0; ay < 4; ay++) {
        int key = 0;
        for (int d = 0; d < 4; d++) {
          const int s = kDir[d][0] + kDir[d][1];  // the one non-zero component
          const int st = kDir[d][0] ? ax : ay;
          if (s < 0 && !(st == 0 || st == 2)) continue;
          if (s > 0 && !(st == 0 || st == 1)) continue;
          key |= 1 << d;
        }
        if (offset[key] < 0) add(key);
      }
  }
  void add(int key) {
    offset[key] = (int)table.size();
    int m = __builtin_popcount(key);
    for (int c = 0; c < (1 << m); c++) table.push_back(link_type(key, c));
  }
};

const Lut& lut() {
  static Lut l;
  return l;
}

// --------------------------------------------------------------- grid ----
// Directions in use for every grid built after oracle_set_adjacency.
int g_dmask = (1 << kNbr) - 1;
// Persistence threshold of the 2-saddle triplets (oracle_set_persistence).
double g_thr = 0.0;

struct Grid {
  int W, H, D;
  int64_t N;
  // directions in use: the 14-neighbour stencil (D1), or with the CPU path's
  // plane adjacency (maxNeighbors = 6 over the table of src/ExtreM.cpp:16-21
  // on a depth-1 field) only +-x and +-y
  int dmask = g_dmask;
  // Existing neighbours of v in direction order: computeAdjacency,
  // src/ExtreM.cpp:543-575 (with the 14 directions, D1).  Returns the count;
  // dirs[] receives the direction index of each neighbour.
  int nbrs(int v, int* out, int* dirs = nullptr) const {
    int x = v % W, y = (v / W) % H, z = v / (W * H);
    int m = 0;
    for (int d = 0; d < kNbr; d++) {
      if (!((dmask >> d) & 1)) continue;
      int nx = x + kDir[d][0], ny = y + kDir[d][1], nz = z + kDir[d][2];
      if (nx < 0 || nx >= W || ny < 0 || ny >= H || nz < 0 || nz >= D)
        continue;
      out[m] = nx + W * (ny + H * nz);
      if (dirs) dirs[m] = d;
      m++;
    }
    return m;
  }
};

// src/ExtreM.cpp:577-583
inline bool isless(int v, int u, const double* f) {
  return f[v] < f[u] || (std::abs(f[v] - f[u]) == 0 && v < u);
}
inline bool islarger(int v, int u, const double* f) {
  return f[v] > f[u] || (std::abs(f[v] - f[u]) == 0 && v > u);
}

// Per-vertex classification.  Order and tie-break follow
// classifyVertex_CUDA (src/ExtreM_cuda.cu:3796-3853); the lower star is the
// direction bitmask of classifyVertex1's lowerStars list
// (src/ExtreM.cpp:723-750) (D9).
int classify_vertex(const Grid& g, const double* f, int v, uint16_t* lower) {
  int nb[kNbr], dr[kNbr];
  int m = g.nbrs(v, nb, dr);
  double c = f[v];
  int key = 0, cfg = 0, lc = 0, uc = 0;
  uint16_t lm = 0;
  for (int k = 0; k < m; k++) {
    int r = nb[k];
    key |= 1 << dr[k];
    if (f[r] < c || (f[r] == c && r < v)) {
      lm |= (uint16_t)(1u << dr[k]);
      lc++;
    } else if (f[r] > c || (f[r] == c && r > v)) {
      cfg |= 1 << k;
      uc++;
    }
  }
  *lower = lm;
  // The GPU path's shortcuts (src/ExtreM_cuda.cu:3846-3853).  The plane
  // adjacency keeps classifyVertex1's rule (src/ExtreM.cpp:880-886) for every
  // configuration: its link may be disconnected, and a vertex with no lower
  // neighbour but two upper link components is a 2-saddle there.
  if (g.dmask == (1 << kNbr) - 1) {
    if (uc == 0) return 4;
    if (lc == 0) return 0;
  }
  const Lut& L = lut();
  return L.table[L.offset[key] + cfg];
}

// ------------------------------------------------------------- engine ----
struct Triplet {
  std::vector<int> roots;  // [0..43] of std::array<int,46>, deduped (D10)
  int saddle;              // [45]
  int first() const { return roots.empty() ? NONE : roots.front(); }  // D11
  int last() const { return roots.empty() ? NONE : roots.back(); }
};

struct CP {
  std::vector<int> saddles2, saddles1, maximum, minimum;
  std::vector<Triplet> trip2, trip1;
  std::unordered_map<int, int> largest, smallest;  // vertex -> list index
};

struct Stats {
  int c_loop_iters, outer_iters, r_loop_calls, s_loop_calls,
      saddle_loop_calls, c_loops_calls;
  int64_t edited;
  int n_max, n_min, n_saddle1, n_saddle2;
  int status;
};

struct Engine {
  Grid g;
  const double* orig;  // input_data
  double* decp;        // decp_data
  double bound;        // bound (xi)
  int max_c_iters, max_outer;
  int step_q = 0;  // 0: the CPU path's rules; q > 0: fixed step -xi/q, threshold q
  bool f32 = false;  // the float32 fork's rule set (src/ExtreM_cuda_f.cu)
  double thr = g_thr;  // persistence filter of the 2-saddle triplets (0: none)
  std::vector<double> delta;       // d_deltaBuffer
  std::vector<uint8_t> counter;    // delta_counter
  std::vector<uint8_t> vt, dvt;    // vertex_type, dec_vertex_type
  std::vector<uint16_t> lstar, dlstar;
  std::vector<int> all_max, all_min, all_saddle;
  std::vector<int> DS, AS, dDS, dAS;
  CP ocp, dcp;
  std::vector<int> or_max_map, or_min_map;  // 14 slots per vertex (D3)
  std::vector<uint8_t> wn, wn_ds;           // wrong_neighbors(_ds)
  std::vector<int> wn_idx, wn_ds_idx;
  int nfc = 0, nfc1 = 0;  // number_of_false_cases(1)
  int wmc = 0, wmc2 = 0, wsc = 0, wminc = 0, wminc2 = 0, wscj = 0;
  std::vector<int> wrmi, wrmi2, wrsi, wrmini, wrmini2, wrsji;  // pair lists
  std::vector<uint8_t> flag_a, flag_b;
  Stats st{};
  bool failed = false;

  // Test instrumentation (tests/test_known_answers.py).  trace, when set,
  // receives every value apply_delta (or the pre-snap) writes, tagged with
  // the loop that proposed it.  mutation != 0 plants one plausible mistake
  // in one rule, so a known-answer test can show that it tells the rule
  // from that alternative; the correction itself always runs with 0.
  struct Edit {
    int phase, vertex;
    double before, after;
  };
  enum Phase { P_PRESNAP = 0, P_C_LOOPS, P_R_LOOPS, P_S_LOOPS, P_S_LOOPS_JOIN, P_SADDLE,
               P_SADDLE_JOIN };
  enum Mutation {
    M_NONE = 0,
    M_WN_ANY_LABEL,      // wrong neighbours: label compared with any original label
    M_WN_ALL,            // wrong neighbours: every mismatch flagged, no return
    M_WN_DESC_ORDER,     // wrong neighbours: saddles in descending id order
    M_FIXPATH_TRUE,      // fixpath direction 0 lowers the true (original) index
    M_FIXPATH_FALSE,     // fixpath direction 1 lowers the false (decp) index
    M_WI_LAST_WINS,      // wrong index max/min: the last saddle per extremum wins
    M_SLOOPS_SWAP,       // s_loops / s_loops_join: the two fix directions swapped
    M_SADDLE_NO_GUARD,   // fix_wrong_index_saddle without its early return
    M_SADDLE_JOIN_LOWER  // fix_wrong_index_saddle direction 1 lowers the true index
  };
  std::vector<Edit>* trace = nullptr;
  int phase = P_PRESNAP;
  int mutation = M_NONE;
  bool apply_enabled = true;  // false: apply_delta leaves decp as it is (rule probes)

  void classify_all(const double* f, uint8_t* types, uint16_t* ls) {
    // classifyAllVertices, src/ExtreM.cpp:1088-1097
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < g.N; i++)
      types[i] = (uint8_t)classify_vertex(g, f, (int)i, &ls[i]);
  }

  // init_delta, src/ExtreM.cpp:1099-1105
  void init_delta() { std::fill(delta.begin(), delta.end(), -4.0 * bound); }

  // swap/atomicCASDouble sequentially: a proposal wins iff larger (D5),
  // src/ExtreM.cpp:1120-1150 with the caller's `d > oldValue` gate.
  void propose(int v, double d) {
    if (step_q) d = -bound / step_q;  // fixed-step rule: every proposal is -xi/q
    // float32 fork: every fix proposal is -bound / (1 << 2) in float
    // (src/ExtreM_cuda_f.cu:2703, 2749, 2804)
    if (f32) d = (double)(-(float)bound / 4.0f);
    if (d > delta[v]) delta[v] = d;
  }
  // get_false_stars' (orig - xi) - decp; the fixed-step rule proposes its
  // step there too, the float32 fork keeps the value, in float
  // (src/ExtreM_cuda_f.cu:2583-2600)
  void propose_star(int v) {
    double d = (orig[v] - bound) - decp[v];
    if (step_q) d = -bound / step_q;
    if (f32) d = (double)(((float)orig[v] - (float)bound) - (float)decp[v]);
    if (d > delta[v]) delta[v] = d;
  }
  double snap(int v) const {  // orig - xi (in float with the float32 fork)
    return f32 ? (double)((float)orig[v] - (float)bound) : orig[v] - bound;
  }
  double lower_mid(int v) const {  // ((orig-xi)+decp)/2 - decp
    return ((orig[v] - bound) + decp[v]) / 2.0 - decp[v];
  }
  double raise_mid(int v) const {  // ((orig+xi)+decp)/2 - decp
    return ((orig[v] + bound) + decp[v]) / 2.0 - decp[v];
  }

  // c_loop, split branch, src/ExtreM.cpp:1152-1236
  void c_loop(int index, int direction) {
    int nb[kNbr];
    int m = g.nbrs(index, nb);
    if (direction == 0) {
      if (vt[index] != 4) {
        propose(index, lower_mid(index));
        return;
      }
      int largest = index;
      for (int k = 0; k < m; k++)
        if (islarger(nb[k], largest, decp)) largest = nb[k];
      if (decp[index] > decp[largest] ||
          (decp[index] == decp[largest] && index > largest))
        return;
      propose(largest, lower_mid(largest));
    } else {
      if (vt[index] != 0) {
        int smallest = index;
        for (int k = 0; k < m; k++)
          if (isless(nb[k], smallest, orig)) smallest = nb[k];
        double d = lower_mid(smallest);
        if (decp[index] > decp[smallest] ||
            (decp[index] == decp[smallest] && index > smallest))
          return;
        propose(smallest, d);
      } else {
        propose(index, lower_mid(index));
      }
    }
  }

  // get_false_criticle_points, src/ExtreM.cpp:1324-1383
  void get_false_critical_points() {
    all_max.clear();
    all_min.clear();
    all_saddle.clear();
    for (int64_t ii = 0; ii < g.N; ii++) {
      int i = (int)ii;
      int t1 = dvt[i], t0 = vt[i];
      if (t1 != t0) {
        if ((t1 == 4) != (t0 == 4)) all_max.push_back(i);
        if ((t1 == 0) != (t0 == 0)) all_min.push_back(i);
        if ((t1 == 2) != (t0 == 2)) all_saddle.push_back(i);
        if ((t1 == 1) != (t0 == 1)) all_saddle.push_back(i);
        if ((t1 == 3) != (t0 == 3)) all_saddle.push_back(i);
      } else if (t1 == 1 || t1 == 2 || t1 == 3) {
        // lower and upper star lists partition the neighbours, so comparing
        // the lower masks compares both lists (src/ExtreM.cpp:1369-1374).
        if (dlstar[i] != lstar[i]) all_saddle.push_back(i);
      }
    }
  }

  // get_false_stars, split branch, src/ExtreM.cpp:1412-1445 (D9)
  void get_false_stars(int index) {
    uint16_t dl = dlstar[index], ol = lstar[index];
    if (dl == ol) return;
    int nb[kNbr], dr[kNbr];
    int m = g.nbrs(index, nb, dr);
    // findDifferences(lowerStar, o_lowerStar): original-lower, dec-upper
    for (int k = 0; k < m; k++)
      if ((ol >> dr[k] & 1) && !(dl >> dr[k] & 1)) propose_star(nb[k]);
    // findDifferences(o_lowerStar, lowerStar): dec-lower, original-upper
    if (dl & ~ol) propose_star(index);
  }

  // apply_delta, split, src/ExtreM.cpp:1486-1503 (D8); with the fixed-step
  // rule, applyDeltaBuffer's edit_type 0 branch (src/ExtreM_cuda.cu:3065-3076)
  // with threshold q.
  void apply_delta() {
    if (!apply_enabled) return;
    const double sentinel = -4.0 * bound;
    for (int64_t t = 0; t < g.N; t++) {
      double dl = delta[t];
      if (dl == sentinel) continue;
      if (trace) trace->push_back({phase, (int)t, decp[t], 0.0});
      if (f32) {  // applyDeltaBuffer, src/ExtreM_cuda_f.cu:2854-2866
        const float df = (float)dl, xf = (float)decp[t], of = (float)orig[t], bf = (float)bound;
        if (std::fabs(df) > 1e-6 && counter[t] < 6 && std::fabs(of - (xf + df)) < bf) {
          decp[t] = (double)(xf + df);
          counter[t] += 1;
        } else {
          counter[t] = 7;
          decp[t] = (double)(of - bf);
        }
        if (trace) trace->back().after = decp[t];
        continue;
      }
      if (step_q) {
        if (counter[t] < step_q - 1 && std::abs(orig[t] - (decp[t] + dl)) <= bound) {
          decp[t] += dl;
          counter[t] += 1;
        } else {
          counter[t] = (uint8_t)step_q;
          decp[t] = orig[t] - bound;
        }
        if (trace) trace->back().after = decp[t];
        continue;
      }
      if (std::abs(dl) > 1e-15 && counter[t] < 5 &&
          std::abs(orig[t] - decp[t] - dl) <= bound) {
        decp[t] += dl;
        counter[t] += 1;
      } else {
        decp[t] = orig[t] - bound;
        counter[t] = 6;
      }
      if (trace) trace->back().after = decp[t];
    }
  }

  // c_loops, src/ExtreM.cpp:1505-1583
  void c_loops() {
    phase = P_C_LOOPS;
    st.c_loops_calls++;
    classify_all(decp, dvt.data(), dlstar.data());
    get_false_critical_points();
    while (!all_max.empty() || !all_min.empty() || !all_saddle.empty()) {
      if (++st.c_loop_iters > max_c_iters) {
        failed = true;
        return;
      }
      init_delta();
      for (int v : all_max) c_loop(v, 0);
      for (int v : all_min) c_loop(v, 1);
      for (int v : all_saddle) get_false_stars(v);
      apply_delta();
      classify_all(decp, dvt.data(), dlstar.data());
      get_false_critical_points();
    }
  }

  // Steepest ascent (descending manifold) labels, src/ExtreM.cpp:144-183;
  // steepest descent (ascending manifold) labels, src/ExtreM.cpp:502-541.
  // Fully compressed roots for every vertex (D7).
  void manifold(const double* f, std::vector<int>& lab, bool ascent) {
    lab.resize(g.N);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < g.N; ii++) {
      int i = (int)ii, nb[kNbr];
      int m = g.nbrs(i, nb);
      int best = i;
      for (int k = 0; k < m; k++) {
        int n = nb[k];
        if (ascent ? (f[best] < f[n] || (f[best] == f[n] && best < n))
                   : (f[best] > f[n] || (f[best] == f[n] && best > n)))
          best = n;
      }
      lab[i] = best;
    }
    // resolve roots: equal to repeated pointer jumping to the fixed point
    std::vector<int> path;
    std::vector<uint8_t> done(g.N, 0);
    for (int64_t ii = 0; ii < g.N; ii++) {
      int v = (int)ii;
      path.clear();
      while (!done[v] && lab[v] != v) {
        path.push_back(v);
        v = lab[v];
      }
      int root = done[v] ? lab[v] : v;
      done[v] = 1;
      for (int p : path) {
        lab[p] = root;
        done[p] = 1;
      }
    }
  }

  // computeMaxLabel / computeMinLabel, src/ExtreM.cpp:2251-2307
  int compute_max_label(int i, const double* f) const {
    int cur = i, nb[kNbr];
    while (true) {
      int next = cur;
      int m = g.nbrs(cur, nb);
      for (int k = 0; k < m; k++)
        if (f[next] < f[nb[k]] || (f[next] == f[nb[k]] && next < nb[k]))
          next = nb[k];
      if (next == cur) return cur;
      cur = next;
    }
  }
  int compute_min_label(int i, const double* f) const {
    int cur = i, nb[kNbr];
    while (true) {
      int next = cur;
      int m = g.nbrs(cur, nb);
      for (int k = 0; k < m; k++)
        if (f[next] > f[nb[k]] || (f[next] == f[nb[k]] && next > nb[k]))
          next = nb[k];
      if (next == cur) return cur;
      cur = next;
    }
  }

  // computeCP + findAscPaths/findDesPaths + sortAndRemoveDuplicates,
  // src/ExtreM.cpp:1920-2040, 1597-1679, 585-595 (D4, D10).
  void compute_cp(CP& cp, const double* f, const std::vector<int>& ds,
                  const std::vector<int>& as, const uint8_t* types) {
    cp.saddles2.clear();
    cp.saddles1.clear();
    cp.maximum.clear();
    cp.minimum.clear();
    for (int64_t ii = 0; ii < g.N; ii++) {
      int v = (int)ii, t = types[v];
      if (t == 2 || t == 3) cp.saddles2.push_back(v);
      if (t == 1 || t == 3) cp.saddles1.push_back(v);
      if (t == 4) cp.maximum.push_back(v);
      if (t == 0) cp.minimum.push_back(v);
    }
    auto asc = [f](int v, int u) {
      return f[v] < f[u] || (std::abs(f[v] - f[u]) == 0 && v < u);
    };
    auto desc = [f](int v, int u) {
      return f[v] > f[u] || (std::abs(f[v] - f[u]) == 0 && v > u);
    };
    std::sort(cp.saddles2.begin(), cp.saddles2.end(), desc);
    std::sort(cp.saddles1.begin(), cp.saddles1.end(), asc);
    std::sort(cp.maximum.begin(), cp.maximum.end(), asc);
    std::sort(cp.minimum.begin(), cp.minimum.end(), desc);
    int nb[kNbr];
    cp.trip2.assign(cp.saddles2.size(), Triplet());
    for (size_t i = 0; i < cp.saddles2.size(); i++) {
      int s = cp.saddles2[i];
      Triplet& t = cp.trip2[i];
      int m = g.nbrs(s, nb);
      // the GPU fork's persistence filter (findAscPaths,
      // src/ExtreM_cuda.cu:1877-1881): a maximum closer than thr stays out
      for (int k = 0; k < m; k++)
        if (islarger(nb[k], s, f) && (thr <= 0.0 || std::fabs(f[ds[nb[k]]] - f[s]) >= thr))
          t.roots.push_back(ds[nb[k]]);
      // rank ascending over a value-ascending maximum list == (value,id) asc
      std::sort(t.roots.begin(), t.roots.end(), asc);
      t.roots.erase(std::unique(t.roots.begin(), t.roots.end()), t.roots.end());
      t.saddle = s;
    }
    cp.trip1.assign(cp.saddles1.size(), Triplet());
    for (size_t i = 0; i < cp.saddles1.size(); i++) {
      int s = cp.saddles1[i];
      Triplet& t = cp.trip1[i];
      int m = g.nbrs(s, nb);
      for (int k = 0; k < m; k++)
        if (isless(nb[k], s, f)) t.roots.push_back(as[nb[k]]);
      // rank ascending over a value-descending minimum list
      std::sort(t.roots.begin(), t.roots.end(), desc);
      t.roots.erase(std::unique(t.roots.begin(), t.roots.end()), t.roots.end());
      t.saddle = s;
    }
  }

  // computelargestSaddlesForMax / computesmallestSaddlesForMin,
  // src/ExtreM.cpp:2309-2371.  trips are in list (value) order here.
  static void extremum_saddle_map(int nExt, const std::vector<Triplet>& trips,
                                  std::unordered_map<int, int>& mp,
                                  const std::vector<int>& ext) {
    mp.clear();
    for (int i = 0; i < nExt && i < (int)ext.size(); i++)
      mp[ext[i]] = (int)trips.size();
    int global = (nExt >= 1 && nExt <= (int)ext.size()) ? ext[nExt - 1] : NONE;
    for (int i = 0; i < (int)trips.size(); i++)
      for (int r : trips[i].roots) {
        if (r == global) continue;
        auto it = mp.find(r);
        if (it == mp.end()) continue;  // D10: root outside the list
        if (i < it->second) it->second = i;
      }
  }

  static void sort_by_saddle(std::vector<Triplet>& t) {
    std::sort(t.begin(), t.end(),
              [](const Triplet& a, const Triplet& b) { return a.saddle < b.saddle; });
  }

  // compute_Max_for_Saddle / compute_Min_for_Saddle, src/ExtreM.cpp:2373-2403
  void compute_saddle_maps() {
    or_max_map.assign((size_t)g.N * kNbr, NONE);
    or_min_map.assign((size_t)g.N * kNbr, NONE);
    int nb[kNbr];
    for (const Triplet& t : ocp.trip2) {
      int s = t.saddle, c = 0, m = g.nbrs(s, nb);
      for (int k = 0; k < m; k++)
        if (islarger(nb[k], s, orig))
          or_max_map[(size_t)s * kNbr + c++] = compute_max_label(nb[k], orig);
    }
    for (const Triplet& t : ocp.trip1) {
      int s = t.saddle, c = 0, m = g.nbrs(s, nb);
      for (int k = 0; k < m; k++)
        if (isless(nb[k], s, orig))
          or_min_map[(size_t)s * kNbr + c++] = compute_min_label(nb[k], orig);
    }
  }

  // get_wrong_split_neighbors / get_wrong_join_neighbors,
  // src/ExtreM.cpp:2405-2450, called in saddle-id order main.cpp:313-323 (D6)
  void wrong_neighbors() {
    nfc = nfc1 = 0;
    std::fill(wn.begin(), wn.end(), 0);
    std::fill(wn_ds.begin(), wn_ds.end(), 0);
    wn_idx.clear();
    wn_ds_idx.clear();
    int nb[kNbr];
    // mismatch of the c-th side label (positional, label_count); with
    // M_WN_ANY_LABEL, a label found nowhere among the original ones
    auto differs = [&](const std::vector<int>& map, int s, int c, int l) {
      if (mutation != M_WN_ANY_LABEL) return l != map[(size_t)s * kNbr + c];
      for (int k = 0; k < kNbr; k++)
        if (map[(size_t)s * kNbr + k] == l) return false;
      return true;
    };
    const bool all = mutation == M_WN_ALL;
    const int n2 = (int)ocp.trip2.size(), n1 = (int)ocp.trip1.size();
    for (int q = 0; q < n2; q++) {
      const Triplet& t = ocp.trip2[mutation == M_WN_DESC_ORDER ? n2 - 1 - q : q];
      int s = t.saddle, c = 0, m = g.nbrs(s, nb);
      for (int k = 0; k < m; k++) {
        int n = nb[k];
        if (!islarger(n, s, decp)) continue;
        int l = compute_max_label(n, decp);
        if (differs(or_max_map, s, c, l) && !wn[n]) {
          wn[n] = 1;
          wn_idx.push_back(n);
          nfc++;
          if (!all) break;
        }
        c++;
      }
    }
    for (int q = 0; q < n1; q++) {
      const Triplet& t = ocp.trip1[mutation == M_WN_DESC_ORDER ? n1 - 1 - q : q];
      int s = t.saddle, c = 0, m = g.nbrs(s, nb);
      for (int k = 0; k < m; k++) {
        int n = nb[k];
        if (!isless(n, s, decp)) continue;
        int l = compute_min_label(n, decp);
        if (differs(or_min_map, s, c, l) && !wn_ds[n]) {
          wn_ds[n] = 1;
          wn_ds_idx.push_back(n);
          nfc1++;
          if (!all) break;
        }
        c++;
      }
    }
  }

  // fixpath, split, src/ExtreM.cpp:2452-2575
  void fixpath(int i, int direction) {
    int true_index = NONE, false_index = NONE;
    int cur = i, nb[kNbr];
    while (true) {
      int nxt = cur, dnxt = cur;
      int m = g.nbrs(cur, nb);
      for (int k = 0; k < m; k++) {
        int n = nb[k];
        if (direction == 0) {
          if (orig[nxt] < orig[n] || (orig[nxt] == orig[n] && nxt < n)) nxt = n;
          if (decp[dnxt] < decp[n] || (decp[dnxt] == decp[n] && dnxt < n)) dnxt = n;
        } else {
          if (orig[nxt] > orig[n] || (orig[nxt] == orig[n] && nxt > n)) nxt = n;
          if (decp[dnxt] > decp[n] || (decp[dnxt] == decp[n] && dnxt > n)) dnxt = n;
        }
      }
      if (nxt != dnxt) {
        true_index = nxt;
        false_index = dnxt;
        break;
      }
      if (nxt == cur) break;
      cur = nxt;
    }
    if (false_index == true_index) return;
    if (mutation == M_FIXPATH_TRUE && direction == 0) std::swap(false_index, true_index);
    if (mutation == M_FIXPATH_FALSE && direction == 1) std::swap(false_index, true_index);
    if (direction == 0)
      propose(false_index, lower_mid(false_index));
    else
      propose(true_index, lower_mid(true_index));
  }

  // r_loops, src/ExtreM.cpp:2577-2593
  void r_loops() {
    phase = P_R_LOOPS;
    st.r_loop_calls++;
    init_delta();
    for (int id : wn_idx) fixpath(id, 0);
    for (int id : wn_ds_idx) fixpath(id, 1);
    apply_delta();
  }

  // M_WI_LAST_WINS: a later saddle replaces the pair recorded for extremum id
  static void overwrite(std::vector<int>& pairs, int id, int false_id) {
    for (size_t k = 0; k + 1 < pairs.size(); k += 2)
      if (pairs[k] == id) pairs[k + 1] = false_id;
  }

  // get_wrong_index_max, src/ExtreM.cpp:2595-2630 (D11)
  void get_wrong_index_max() {
    wmc = wmc2 = 0;
    wrmi.clear();
    wrmi2.clear();
    std::fill(flag_a.begin(), flag_a.end(), 0);
    std::fill(flag_b.begin(), flag_b.end(), 0);
    size_t n = std::min(dcp.trip2.size(), ocp.trip2.size());
    for (size_t i = 0; i < n; i++) {
      const Triplet &o = ocp.trip2[i], &d = dcp.trip2[i];
      if (o.first() != d.first() && o.first() != NONE && d.first() != NONE) {  // D16
        int id = o.first();
        if (id != NONE && !flag_a[id]) {
          wrmi.push_back(o.first());
          wrmi.push_back(d.first());
          flag_a[id] = 1;
          wmc++;
        } else if (id != NONE && mutation == M_WI_LAST_WINS) {
          overwrite(wrmi, id, d.first());
        }
      }
      if (o.last() != d.last() && o.last() != NONE && d.last() != NONE) {  // D16
        int id = o.last();
        if (id != NONE && !flag_b[id]) {
          wrmi2.push_back(o.last());
          wrmi2.push_back(d.last());
          flag_b[id] = 1;
          wmc2++;
        } else if (id != NONE && mutation == M_WI_LAST_WINS) {
          overwrite(wrmi2, id, d.last());
        }
      }
    }
  }

  // get_wrong_index_min, src/ExtreM.cpp:2632-2670
  void get_wrong_index_min() {
    wminc = wminc2 = 0;
    wrmini.clear();
    wrmini2.clear();
    std::fill(flag_a.begin(), flag_a.end(), 0);
    std::fill(flag_b.begin(), flag_b.end(), 0);
    size_t n = std::min(dcp.trip1.size(), ocp.trip1.size());
    for (size_t i = 0; i < n; i++) {
      const Triplet &o = ocp.trip1[i], &d = dcp.trip1[i];
      if (o.first() != d.first() && o.first() != NONE && d.first() != NONE) {  // D16
        int id = o.first();
        if (id != NONE && !flag_b[id]) {
          wrmini2.push_back(o.first());
          wrmini2.push_back(d.first());
          flag_b[id] = 1;
          wminc2++;
        } else if (id != NONE && mutation == M_WI_LAST_WINS) {
          overwrite(wrmini2, id, d.first());
        }
      }
      if (o.last() != d.last() && o.last() != NONE && d.last() != NONE) {  // D16
        int id = o.last();
        if (id != NONE && !flag_a[id]) {
          wrmini.push_back(o.last());
          wrmini.push_back(d.last());
          flag_a[id] = 1;
          wminc++;
        } else if (id != NONE && mutation == M_WI_LAST_WINS) {
          overwrite(wrmini, id, d.last());
        }
      }
    }
  }

  // get_wrong_index_saddles(_join), src/ExtreM.cpp:2672-2715 (D12)
  int wrong_index_saddles(const std::vector<int>& dext,
                          const std::unordered_map<int, int>& dmap,
                          std::unordered_map<int, int>& omap,
                          const std::vector<int>& dsad,
                          const std::vector<int>& osad, std::vector<int>& out) {
    out.clear();
    int cnt = 0;
    for (int ext : dext) {
      auto it = dmap.find(ext);
      if (it == dmap.end()) continue;
      int di = it->second;
      int oi = omap[ext];  // operator[] as the reference
      int dv = (di >= 0 && di < (int)dsad.size()) ? dsad[di] : NONE;
      int ov = (oi >= 0 && oi < (int)osad.size()) ? osad[oi] : NONE;
      if (dv != ov && dv != NONE && ov != NONE) {  // D12
        out.push_back(ov);
        out.push_back(dv);
        cnt++;
      }
    }
    return cnt;
  }

  // fix_wrong_index_saddle, src/ExtreM.cpp:2811-2836 (D13)
  void fix_wrong_index_saddle(int t, int f, int direction) {
    if (t == NONE || f == NONE) return;
    if (mutation != M_SADDLE_NO_GUARD && (decp[t] > decp[f] || (decp[t] == decp[f] && t > f)))
      return;
    if (direction == 0)
      propose(f, lower_mid(f));
    else if (mutation == M_SADDLE_JOIN_LOWER)
      propose(t, lower_mid(t));
    else
      propose(t, raise_mid(t));
  }

  // s_loops, src/ExtreM.cpp:2838-2862
  void s_loops() {
    phase = P_S_LOOPS;
    st.s_loop_calls++;
    get_wrong_index_max();
    init_delta();
    const int a = mutation == M_SLOOPS_SWAP ? 1 : 0;  // dir 0: the true index
    for (int i = 0; i < wmc; i++) {
      int v = wrmi[2 * i + a];
      if (v != NONE) propose(v, lower_mid(v));
    }
    for (int i = 0; i < wmc2; i++) {  // dir 1: the false index
      int v = wrmi2[2 * i + 1 - a];
      if (v != NONE) propose(v, lower_mid(v));
    }
    apply_delta();
  }
  // s_loops_join, src/ExtreM.cpp:2864-2886
  void s_loops_join() {
    phase = P_S_LOOPS_JOIN;
    st.s_loop_calls++;
    get_wrong_index_min();
    init_delta();
    const int a = mutation == M_SLOOPS_SWAP ? 1 : 0;  // dir 0: the true index
    for (int i = 0; i < wminc; i++) {
      int v = wrmini[2 * i + a];
      if (v != NONE) propose(v, lower_mid(v));
    }
    for (int i = 0; i < wminc2; i++) {  // dir 1: the false index
      int v = wrmini2[2 * i + 1 - a];
      if (v != NONE) propose(v, lower_mid(v));
    }
    apply_delta();
  }
  // saddle_loops / saddle_loops_join, src/ExtreM.cpp:2888-2916
  void saddle_loops() {
    phase = P_SADDLE;
    st.saddle_loop_calls++;
    wsc = wrong_index_saddles(dcp.maximum, dcp.largest, ocp.largest,
                              dcp.saddles2, ocp.saddles2, wrsi);
    init_delta();
    for (int i = 0; i < wsc; i++) fix_wrong_index_saddle(wrsi[2 * i], wrsi[2 * i + 1], 0);
    apply_delta();
  }
  void saddle_loops_join() {
    phase = P_SADDLE_JOIN;
    st.saddle_loop_calls++;
    wscj = wrong_index_saddles(dcp.minimum, dcp.smallest, ocp.smallest,
                               dcp.saddles1, ocp.saddles1, wrsji);
    init_delta();
    for (int i = 0; i < wscj; i++)
      fix_wrong_index_saddle(wrsji[2 * i], wrsji[2 * i + 1], 1);
    apply_delta();
  }

  // dec manifolds + computeCP + largest/smallest + sort by saddle id
  // (main.cpp:290-304, repeated at 342-355 and 393-408)
  void dec_cp_block() {
    manifold(decp, dDS, true);
    manifold(decp, dAS, false);
    compute_cp(dcp, decp, dDS, dAS, dvt.data());
    extremum_saddle_map(nMax, dcp.trip2, dcp.largest, dcp.maximum);
    extremum_saddle_map(nMin, dcp.trip1, dcp.smallest, dcp.minimum);
    sort_by_saddle(dcp.trip2);
    sort_by_saddle(dcp.trip1);
  }
  int nMax = 0, nMin = 0;

  bool any_counter() const {
    return nfc > 0 || nfc1 > 0 || !all_max.empty() || !all_min.empty() ||
           !all_saddle.empty() || wmc > 0 || wmc2 > 0 || wsc > 0 ||
           wminc > 0 || wminc2 > 0 || wscj > 0;
  }

  // main.cpp:65-446
  int run() {
    const int64_t N = g.N;
    delta.assign(N, -4.0 * bound);
    counter.assign(N, 0);
    vt.assign(N, 5);
    dvt.assign(N, 5);
    lstar.assign(N, 0);
    dlstar.assign(N, 0);
    wn.assign(N, 0);
    wn_ds.assign(N, 0);
    flag_a.assign(N, 0);
    flag_b.assign(N, 0);

    // main.cpp:67-75: global min / max of orig by (value, id), pre-snap
    int gmin = 0, gmax = 0;
    for (int64_t ii = 1; ii < N; ii++) {
      int i = (int)ii;
      if (isless(i, gmin, orig)) gmin = i;
      if (islarger(i, gmax, orig)) gmax = i;
    }
    if (trace) trace->push_back({P_PRESNAP, gmin, decp[gmin], snap(gmin)});
    decp[gmin] = snap(gmin);
    if (trace) trace->push_back({P_PRESNAP, gmax, decp[gmax], snap(gmax)});
    decp[gmax] = snap(gmax);

    // main.cpp:85-89
    classify_all(orig, vt.data(), lstar.data());
    manifold(orig, DS, true);
    manifold(orig, AS, false);

    // main.cpp:100-119
    compute_cp(ocp, orig, DS, AS, vt.data());
    nMax = (int)ocp.maximum.size();
    nMin = (int)ocp.minimum.size();
    st.n_max = nMax;
    st.n_min = nMin;
    st.n_saddle2 = (int)ocp.saddles2.size();
    st.n_saddle1 = (int)ocp.saddles1.size();
    extremum_saddle_map(nMax, ocp.trip2, ocp.largest, ocp.maximum);
    extremum_saddle_map(nMin, ocp.trip1, ocp.smallest, ocp.minimum);
    sort_by_saddle(ocp.trip2);
    sort_by_saddle(ocp.trip1);

    // main.cpp:122-131 (preserve_join_tree = true, main.cpp:21)
    for (int s : ocp.saddles2) {
      if (trace) trace->push_back({P_PRESNAP, s, decp[s], snap(s)});
      decp[s] = snap(s);
    }
    for (int s : ocp.saddles1) {
      if (trace) trace->push_back({P_PRESNAP, s, decp[s], snap(s)});
      decp[s] = snap(s);
    }

    c_loops();  // main.cpp:139
    if (failed) return fail();

    compute_saddle_maps();  // main.cpp:195-207

    manifold(decp, dDS, true);  // main.cpp:172-191
    manifold(decp, dAS, false);
    compute_cp(dcp, decp, dDS, dAS, dvt.data());
    sort_by_saddle(dcp.trip2);
    sort_by_saddle(dcp.trip1);

    wrong_neighbors();  // main.cpp:210-226

    if (nfc == 0) {  // main.cpp:230-254
      manifold(decp, dDS, true);
      compute_cp(dcp, decp, dDS, dAS, dvt.data());
      extremum_saddle_map(nMax, dcp.trip2, dcp.largest, dcp.maximum);
      sort_by_saddle(dcp.trip2);
      sort_by_saddle(dcp.trip1);
      s_loops();
      if (wmc == 0) saddle_loops();
    }
    if (nfc1 == 0) {  // main.cpp:255-274
      manifold(decp, dDS, true);
      manifold(decp, dAS, false);
      compute_cp(dcp, decp, dDS, dAS, dvt.data());
      extremum_saddle_map(nMin, dcp.trip1, dcp.smallest, dcp.minimum);
      sort_by_saddle(dcp.trip2);
      sort_by_saddle(dcp.trip1);
      s_loops_join();
      if (wminc == 0) saddle_loops_join();
    }

    while (any_counter()) {  // main.cpp:277-446
      if (++st.outer_iters > max_outer) return fail();
      r_loops();
      for (int rep = 0; rep < 3; rep++) {
        c_loops();
        if (failed) return fail();
        dec_cp_block();
        wrong_neighbors();
        if (rep < 2) {
          if (nfc == 0) {
            s_loops();
            if (wmc == 0) saddle_loops();
          }
          if (nfc1 == 0) {
            s_loops_join();
            if (wminc == 0) saddle_loops_join();
          }
        } else {
          if (nfc == 0) {  // main.cpp:432-437
            get_wrong_index_max();
            wsc = wrong_index_saddles(dcp.maximum, dcp.largest, ocp.largest,
                                      dcp.saddles2, ocp.saddles2, wrsi);
          }
          if (nfc1 == 0) {  // main.cpp:439-444
            get_wrong_index_min();
            wscj = wrong_index_saddles(dcp.minimum, dcp.smallest, ocp.smallest,
                                       dcp.saddles1, ocp.saddles1, wrsji);
          }
        }
      }
    }
    for (int64_t i = 0; i < N; i++) st.edited += counter[i] != 0;
    st.status = 0;
    return 0;
  }
  int fail() {
    st.status = -3;
    return -3;
  }
};

// --------------------------------------------------- split tree check ----
// compute_MergeT of the reference (src/ExtreM.cpp:2930-2952) for one field,
// as main.cpp:454-463 runs it after the correction: computeCP's value-sorted
// lists and triplets (src/ExtreM.cpp:1920-2040, 1597-1636),
// constructPersistencePairs (:2132-2249) and constructMergeTree (:1863-1918).
//   D14 getSortedPositions (:1840-1860) ranks by value with std::sort, which
//       leaves ties unordered; ties are ranked by vertex id, as
//       getSortedIndexes (:1826-1838) does.
//   D15 each field is processed with its own types, labels and value order
//       (main.cpp:459-461 sorts the original field's lists by decp).
struct SplitTree {
  std::vector<int> maximum;                  // ascending (value, id): rank order
  std::vector<std::pair<int, int>> pairs;    // per rank: (saddle vertex, max vertex)
  std::vector<std::pair<int, int>> mtrip;    // per rank: (saddle index, parent rank)
  std::vector<std::vector<int>> branch;      // per rank: vertex ids, descending order
  std::vector<int> saddles2;                 // descending (value, id)
  int global_min = NONE;
  int rounds = 0;
};

SplitTree split_tree(const Grid& g, const double* f) {
  SplitTree t;
  std::vector<uint8_t> type(g.N);
  std::vector<uint16_t> low(g.N);
  for (int64_t i = 0; i < g.N; i++) type[i] = (uint8_t)classify_vertex(g, f, (int)i, &low[i]);
  Engine e;
  e.g = g;
  std::vector<int> DS;
  e.manifold(f, DS, true);
  for (int64_t i = 0; i < g.N; i++) {
    if (type[i] == 2 || type[i] == 3) t.saddles2.push_back((int)i);
    if (type[i] == 4) t.maximum.push_back((int)i);
  }
  std::sort(t.saddles2.begin(), t.saddles2.end(),
            [&](int v, int u) { return islarger(v, u, f); });
  std::sort(t.maximum.begin(), t.maximum.end(), [&](int v, int u) { return isless(v, u, f); });
  const int nMax = (int)t.maximum.size(), nS = (int)t.saddles2.size();
  std::vector<int> rank(g.N, NONE);  // tempArray
  for (int i = 0; i < nMax; i++) rank[t.maximum[i]] = i;
  // findAscPaths: ranks of the maxima reached from the upper neighbours,
  // sorted and deduplicated (an empty triplet stays empty, D11)
  std::vector<std::vector<int>> trip(nS);
  for (int i = 0; i < nS; i++) {
    int v = t.saddles2[i], nb[kNbr];
    int m = g.nbrs(v, nb);
    for (int k = 0; k < m; k++)
      if (islarger(nb[k], v, f)) trip[i].push_back(rank[DS[nb[k]]]);
    std::sort(trip[i].begin(), trip[i].end());
    trip[i].erase(std::unique(trip[i].begin(), trip[i].end()), trip[i].end());
  }
  int gmin = 0;
  for (int64_t i = 1; i < g.N; i++)
    if (isless((int)i, gmin, f)) gmin = (int)i;
  t.global_min = gmin;
  t.pairs.assign(nMax, {0, 0});
  t.mtrip.assign(nMax, {0, 0});
  if (nMax == 0) return t;
  // constructPersistencePairs
  std::vector<int> ptr(nMax);
  std::iota(ptr.begin(), ptr.end(), 0);
  const int globalMax = nMax - 1;
  bool changed = true;
  while (changed) {
    changed = false;
    std::vector<int> largest(nMax, nS);
    for (int i = 0; i < nS; i++)
      for (int m : trip[i])
        if (m != globalMax) largest[m] = std::min(i, largest[m]);
    std::vector<int> active;
    for (int m = 0; m < nMax; m++) {
      int ls = largest[m];
      if (ls < nS && trip[ls].front() == m) {
        changed = true;
        t.pairs[m] = {t.saddles2[ls], t.maximum[m]};
        int lmax = trip[ls].back();
        ptr[m] = lmax;
        t.mtrip[m] = {ls, lmax};
        active.push_back(m);
      }
    }
    while (!active.empty()) {
      std::vector<int> keep;
      for (int v : active) {
        ptr[v] = ptr[ptr[v]];
        if (ptr[v] != ptr[ptr[v]]) keep.push_back(v);
      }
      active.swap(keep);
    }
    for (auto& tr : trip) {
      for (int& r : tr) r = ptr[r];
      std::sort(tr.begin(), tr.end());
      tr.erase(std::unique(tr.begin(), tr.end()), tr.end());
      if (tr.size() == 1) tr.clear();
    }
    t.rounds++;
  }
  t.pairs[nMax - 1] = {gmin, t.maximum[nMax - 1]};
  t.mtrip[nMax - 1] = {gmin, nMax - 1};
  // constructMergeTree: follow larger saddles to the parent
  t.mtrip[nMax - 1].second = nMax - 1;
  bool same = false;
  while (!same) {
    same = true;
    for (int i = 0; i < nMax; i++) {
      auto next = t.mtrip[t.mtrip[i].second];
      if (next.second != t.mtrip[i].second && next.first < t.mtrip[i].first) {
        t.mtrip[i].second = next.second;
        same = false;
      }
    }
  }
  // branches: (order, id) pairs sorted descending; order = rank by (value, id) (D14)
  std::vector<std::vector<std::pair<int, int>>> br(nMax);
  std::vector<int> order_of;  // ranks only of the vertices used
  auto order = [&](int v) {   // rank of v among all vertices by (value, id)
    int64_t r = 0;
    for (int64_t u = 0; u < g.N; u++) r += isless((int)u, v, f);
    return (int)r;
  };
  for (int b = 0; b < nMax; b++) {
    int mx = t.maximum[b];
    br[b].push_back({order(mx), mx});
    int parent = t.mtrip[b].second;
    if (parent != b && !t.saddles2.empty()) {
      int sad = t.saddles2[t.mtrip[b].first];
      int os = order(sad);
      br[b].push_back({os, sad});
      br[parent].push_back({os, sad});
    } else {
      br[b].push_back({-1, t.mtrip[b].first});
    }
  }
  t.branch.resize(nMax);
  for (int b = 0; b < nMax; b++) {
    std::sort(br[b].begin(), br[b].end(), std::greater<>());
    for (auto& pr : br[b]) t.branch[b].push_back(pr.second);
  }
  return t;
}

}  // namespace

extern "C" {

// Split-tree check of one field (compute_MergeT): per maximum in rank order
// (ascending (value, id)): its vertex, the persistence pair's saddle vertex
// (global minimum for the global maximum), the merge-tree parent's vertex,
// and the branch's vertices (descending order) flattened with offsets.
// Returns the number of maxima, or -(needed) if a buffer is too small.
int oracle_split_tree(const double* f, int W, int H, int D, int* max_v, int* pair_s,
                      int* parent_v, int* branch_off, int* branch_v, int cap_max, int cap_branch,
                      int* rounds) {
  Grid g{W, H, D, (int64_t)W * H * D};
  SplitTree t = split_tree(g, f);
  const int nMax = (int)t.maximum.size();
  int total = 0;
  for (auto& b : t.branch) total += (int)b.size();
  if (nMax > cap_max || total > cap_branch) return -std::max(nMax, total);
  int k = 0;
  for (int i = 0; i < nMax; i++) {
    max_v[i] = t.maximum[i];
    pair_s[i] = t.pairs[i].first;
    parent_v[i] = t.maximum[t.mtrip[i].second];
    branch_off[i] = k;
    for (int v : t.branch[i]) branch_v[k++] = v;
  }
  branch_off[nMax] = k;
  *rounds = t.rounds;
  return nMax;
}


// Writes the reference LUT layout (24 800 int32 for the 27 keys of
// src/ExtreM_cuda.cu:84-110) into out; returns the number of entries, or the
// required size if cap is too small.
int oracle_lut(int32_t* out, int cap) {
  const Lut& L = lut();
  if (cap < L.ref_size) return L.ref_size;
  for (int i = 0; i < L.ref_size; i++) out[i] = L.table[i];
  return L.ref_size;
}

// Link-component type for one boundary key and upper config.
int oracle_link_type(int key, int config) { return link_type(key, config); }

// Adjacency of every later call: 0 the 14-neighbour stencil, 1 the CPU
// path's plane adjacency (+-x, +-y; depth-1 fields).
int oracle_set_adjacency(int plane) {
  g_dmask = plane ? 0xF : (1 << kNbr) - 1;
  return 0;
}
// Persistence threshold of the 2-saddle triplets for every later correction
// (0: none).
int oracle_set_persistence(double thr) {
  g_thr = thr;
  return 0;
}

int oracle_classify(const double* f, int W, int H, int D, uint8_t* types,
                    uint16_t* lower) {
  Grid g{W, H, D, (int64_t)W * H * D};
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < g.N; i++)
    types[i] = (uint8_t)classify_vertex(g, f, (int)i, &lower[i]);
  return 0;
}

// Fully compressed steepest-ascent (ds) and steepest-descent (as) roots.
int oracle_labels(const double* f, int W, int H, int D, int32_t* ds,
                  int32_t* as) {
  Engine e;
  e.g = Grid{W, H, D, (int64_t)W * H * D};
  std::vector<int> a, b;
  e.manifold(f, a, true);
  e.manifold(f, b, false);
  std::memcpy(ds, a.data(), sizeof(int) * a.size());
  std::memcpy(as, b.data(), sizeof(int) * b.size());
  return 0;
}

// Path-walk labels (computeMaxLabel / computeMinLabel) for one vertex.
int oracle_max_label(const double* f, int W, int H, int D, int v) {
  Engine e;
  e.g = Grid{W, H, D, (int64_t)W * H * D};
  return e.compute_max_label(v, f);
}
int oracle_min_label(const double* f, int W, int H, int D, int v) {
  Engine e;
  e.g = Grid{W, H, D, (int64_t)W * H * D};
  return e.compute_min_label(v, f);
}

// Full correction (main.cpp:65-446).  decp is edited in place; counts (may be
// null) receives delta_counter; stats (may be null) receives 12 int64:
// status, c_loop_iters, outer_iters, r_loop_calls, s_loop_calls,
// saddle_loop_calls, c_loops_calls, edited, n_max, n_min, n_saddle1,
// n_saddle2.  Returns 0, or -3 when an iteration cap is hit.
int oracle_correct_rule(const double* orig, double* decp, int W, int H, int D, double xi,
                        int max_c_iters, int max_outer, uint8_t* counts, int64_t* stats,
                        int threads, int step_q);
int oracle_correct(const double* orig, double* decp, int W, int H, int D,
                   double xi, int max_c_iters, int max_outer, uint8_t* counts,
                   int64_t* stats, int threads) {
  return oracle_correct_rule(orig, decp, W, H, D, xi, max_c_iters, max_outer, counts, stats,
                             threads, 0);
}

// The same with a rule set: step_q = 0 for the CPU path's midpoint / snap
// rules, q > 0 for the GPU fork's fixed step -xi/q and threshold q
// (edit_type 0: src/ExtreM_cuda.cu:2905 and the other `-bound / (q)` sites,
// applyDeltaBuffer :3065-3076).
// One rule of the outer loop on a given state, for known-answer tests: the
// original field's and decp's classification, labels, CP lists, triplets and
// extremum -> saddle maps are computed as the loop has them (decp with its
// own types), then rule r runs once:
//   0 get_wrong_split/join_neighbors   out: flagged split ids, then join ids
//   1 r_loops' fixpath proposals        out: (vertex, delta) pairs
//   2 get_wrong_index_max + s_loops     out: pairs (wmc), pairs (wmc2), proposals
//   3 get_wrong_index_min + s_loops_join  (as 2)
//   4 get_wrong_index_saddles + saddle_loops' fixes (split)  out: pairs, proposals
//   5 the join versions of 4
// Integer outputs go to iout (counts first: how many of each list), proposals
// to (pv, pd).  Nothing is applied.  Returns 0, or -22.
int oracle_rule_probe(const double* orig, double* decp, int W, int H, int D, double xi, int r,
                      int mutation, int32_t* iout, int icap, int32_t* pv, double* pd,
                      int pcap, int32_t* np) {
  if (W <= 0 || H <= 0 || D <= 0 || !(xi > 0) || r < 0 || r > 5) return -22;
  Engine e;
  e.g = Grid{W, H, D, (int64_t)W * H * D};
  e.orig = orig;
  e.decp = decp;
  e.bound = xi;
  e.mutation = mutation;
  const int64_t N = e.g.N;
  e.delta.assign(N, -4.0 * xi);
  e.counter.assign(N, 0);
  e.vt.assign(N, 5);
  e.dvt.assign(N, 5);
  e.lstar.assign(N, 0);
  e.dlstar.assign(N, 0);
  e.wn.assign(N, 0);
  e.wn_ds.assign(N, 0);
  e.flag_a.assign(N, 0);
  e.flag_b.assign(N, 0);
#ifdef _OPENMP
  const int threads = omp_get_max_threads();
  omp_set_num_threads(1);
#endif
  e.classify_all(orig, e.vt.data(), e.lstar.data());
  e.classify_all(decp, e.dvt.data(), e.dlstar.data());
  e.manifold(orig, e.DS, true);
  e.manifold(orig, e.AS, false);
  e.compute_cp(e.ocp, orig, e.DS, e.AS, e.vt.data());
  e.nMax = (int)e.ocp.maximum.size();
  e.nMin = (int)e.ocp.minimum.size();
  Engine::extremum_saddle_map(e.nMax, e.ocp.trip2, e.ocp.largest, e.ocp.maximum);
  Engine::extremum_saddle_map(e.nMin, e.ocp.trip1, e.ocp.smallest, e.ocp.minimum);
  Engine::sort_by_saddle(e.ocp.trip2);
  Engine::sort_by_saddle(e.ocp.trip1);
  e.compute_saddle_maps();
  e.dec_cp_block();
  std::vector<int> ints;
  auto list = [&](const std::vector<int>& v, int n) {
    ints.push_back(n);
    ints.insert(ints.end(), v.begin(), v.begin() + std::min<size_t>(v.size(), 2 * (size_t)n));
  };
  e.init_delta();
  switch (r) {
    case 0:
    case 1:
      e.wrong_neighbors();
      ints.push_back(e.nfc);
      ints.insert(ints.end(), e.wn_idx.begin(), e.wn_idx.end());
      ints.push_back(e.nfc1);
      ints.insert(ints.end(), e.wn_ds_idx.begin(), e.wn_ds_idx.end());
      if (r == 1) {
        for (int id : e.wn_idx) e.fixpath(id, 0);
        for (int id : e.wn_ds_idx) e.fixpath(id, 1);
      }
      break;
    case 2:
    case 3: {
      // s_loops / s_loops_join without their apply_delta
      e.apply_enabled = false;
      if (r == 2) {
        e.s_loops();
        list(e.wrmi, e.wmc);
        list(e.wrmi2, e.wmc2);
      } else {
        e.s_loops_join();
        list(e.wrmini, e.wminc);
        list(e.wrmini2, e.wminc2);
      }
      break;
    }
    case 4:
    case 5:
      e.apply_enabled = false;
      if (r == 4) {
        e.saddle_loops();
        list(e.wrsi, e.wsc);
      } else {
        e.saddle_loops_join();
        list(e.wrsji, e.wscj);
      }
      break;
  }
#ifdef _OPENMP
  omp_set_num_threads(threads);
#endif
  if ((int)ints.size() > icap) return -22;
  std::copy(ints.begin(), ints.end(), iout);
  int k = 0;
  const double sent = -4.0 * xi;
  for (int64_t t = 0; t < N; t++)
    if (e.delta[t] != sent) {
      if (k >= pcap) return -22;
      pv[k] = (int32_t)t;
      pd[k] = e.delta[t];
      k++;
    }
  *np = k;
  return 0;
}

// The correction with its edit trace (phase, vertex, before, after per
// written value, at most cap entries; returns the number of edits, or -1 -
// needed when cap is too small) and one planted mistake (mutation, 0 = none).
int oracle_correct_trace(const double* orig, double* decp, int W, int H, int D, double xi,
                         int mutation, uint8_t* counts, int64_t* stats, int32_t* phase,
                         int32_t* vertex, double* before, double* after, int cap) {
  if (W <= 0 || H <= 0 || D <= 0 || !(xi > 0)) return -22;
  Engine e;
  std::vector<Engine::Edit> tr;
  e.g = Grid{W, H, D, (int64_t)W * H * D};
  e.orig = orig;
  e.decp = decp;
  e.bound = xi;
  // the known-answer fields are tiny: one thread, and a planted mistake that
  // never converges stops early
  e.max_c_iters = 200;
  e.max_outer = 40;
  e.trace = &tr;
  e.mutation = mutation;
#ifdef _OPENMP
  const int threads = omp_get_max_threads();
  omp_set_num_threads(1);
#endif
  int rc = e.run();
#ifdef _OPENMP
  omp_set_num_threads(threads);
#endif
  if (counts) std::memcpy(counts, e.counter.data(), e.counter.size());
  if (stats) {
    const Stats& s = e.st;
    int64_t v[12] = {rc, s.c_loop_iters, s.outer_iters, s.r_loop_calls,
                     s.s_loop_calls, s.saddle_loop_calls, s.c_loops_calls,
                     s.edited, s.n_max, s.n_min, s.n_saddle1, s.n_saddle2};
    std::memcpy(stats, v, sizeof(v));
  }
  if ((int)tr.size() > cap) return -1 - (int)tr.size();
  for (size_t i = 0; i < tr.size(); i++) {
    phase[i] = tr[i].phase;
    vertex[i] = tr[i].vertex;
    before[i] = tr[i].before;
    after[i] = tr[i].after;
  }
  return (int)tr.size();
}

int oracle_correct_rules(const double* orig, double* decp, int W, int H, int D, double xi,
                         int max_c_iters, int max_outer, uint8_t* counts, int64_t* stats,
                         int threads, int step_q, int f32);
int oracle_correct_rule(const double* orig, double* decp, int W, int H, int D, double xi,
                        int max_c_iters, int max_outer, uint8_t* counts, int64_t* stats,
                        int threads, int step_q) {
  return oracle_correct_rules(orig, decp, W, H, D, xi, max_c_iters, max_outer, counts, stats,
                              threads, step_q, 0);
}
// f32: the float32 fork's rule set (bound, orig and decp in float; inputs
// that float cannot hold exactly are refused).
int oracle_correct_rules(const double* orig, double* decp, int W, int H, int D, double xi,
                         int max_c_iters, int max_outer, uint8_t* counts, int64_t* stats,
                         int threads, int step_q, int f32) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  if (W <= 0 || H <= 0 || D <= 0 || !(xi > 0) || step_q < 0 || step_q > 254) return -22;
  if (f32 && step_q) return -22;
  const int64_t n = (int64_t)W * H * D;
  if (f32) {
    xi = (double)(float)xi;
    for (int64_t i = 0; i < n; i++)
      if ((double)(float)orig[i] != orig[i] || (double)(float)decp[i] != decp[i]) return -22;
  }
  Engine e;
  e.g = Grid{W, H, D, (int64_t)W * H * D};
  e.orig = orig;
  e.decp = decp;
  e.bound = xi;
  e.f32 = f32 != 0;
  e.step_q = step_q;
  e.max_c_iters = max_c_iters > 0 ? max_c_iters : 100000;
  e.max_outer = max_outer > 0 ? max_outer : 10000;
  int rc = e.run();
  if (counts) std::memcpy(counts, e.counter.data(), e.counter.size());
  if (stats) {
    const Stats& s = e.st;
    int64_t v[12] = {rc, s.c_loop_iters, s.outer_iters, s.r_loop_calls,
                     s.s_loop_calls, s.saddle_loop_calls, s.c_loops_calls,
                     s.edited, s.n_max, s.n_min, s.n_saddle1, s.n_saddle2};
    std::memcpy(stats, v, sizeof(v));
  }
  return rc;
}

// 3-bit LSB-first packing of per-vertex edit counts (encodeTo3BitBitmap,
// src/ExtreM.cpp:220-250).  Returns bytes written or -1 for a value > 6.
int64_t oracle_bitmap3(const uint8_t* vals, int64_t n, uint8_t* out) {
  int64_t nb = 0;
  unsigned cur = 0;
  int bit = 0;
  for (int64_t i = 0; i < n; i++) {
    unsigned v = vals[i];
    if (v > 6) return -1;
    cur |= v << bit;
    bit += 3;
    if (bit >= 8) {
      out[nb++] = (uint8_t)cur;
      bit -= 8;
      cur = v >> (3 - bit);
    }
  }
  if (bit > 0) out[nb++] = (uint8_t)cur;
  return nb;
}

// cost() of src/ExtreM.cpp:272-389 without the files: indexs / edits /
// cnt / ratio (:273-294), the 3-bit bitmap of indexs and calculatePSNR
// (:252-270: serial sum of squared differences, -20 log10(sqrt(mse)/range))
// of decp_before and decp_after against orig over max - min of orig.
// out = {changed, snapped, edit_ratio, psnr, fixed_psnr}; bitmap receives
// ceil(3N/8) bytes, edits the snapped edits in index order.
int oracle_cost(const double* orig, const double* before, const double* after,
                const uint8_t* counts, int64_t n, double xi, double* out, uint8_t* bitmap,
                double* edits) {
  std::vector<uint8_t> indexs((size_t)n, 0);
  int64_t cnt = 0, ne = 0;
  for (int64_t i = 0; i < n; i++) {
    if (before[i] != after[i]) {
      indexs[i] = counts[i];
      if (indexs[i] == 6) edits[ne++] = -(before[i] - (orig[i] - xi));
      cnt++;
    }
  }
  if (oracle_bitmap3(indexs.data(), n, bitmap) < 0) return -22;
  double mn = orig[0], mx = orig[0];
  for (int64_t i = 0; i < n; i++) {
    mn = std::min(mn, orig[i]);
    mx = std::max(mx, orig[i]);
  }
  auto psnr = [&](const double* c) {
    double mse = 0.0;
    for (int64_t i = 0; i < n; i++) mse += std::pow(orig[i] - c[i], 2);
    mse /= (double)n;
    if (mse == 0) return std::numeric_limits<double>::infinity();
    return -20.0 * std::log10(std::sqrt(mse) / (mx - mn));
  };
  out[0] = (double)cnt;
  out[1] = (double)ne;
  out[2] = (double)cnt / (double)n;
  out[3] = psnr(before);
  out[4] = psnr(after);
  return 0;
}

}  // extern "C"
