// dropin_main.cpp -- TEST PROGRAM: the reference's correction call written
// against include/api/ExtreM.h and linked with libExtreMz, calling the API in
// the order of the reference's main.cpp:65-507 (set-up, pre-snap, c_loops,
// the wrong-neighbour / wrong-index / saddle passes, the outer loop, the
// split-tree comparison, fixed.bin).
//
//   dropin_main <field.f64> <W> <H> <D> <er> <maxNeighbors> <out.bin>
//
// The lossy field comes from getdata (EXTREMZ_DECP names its file).  Prints
// one line of counters; exit 0 when the loop ended.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ExtreM.h"

namespace {

const std::string kSplit = "split";
constexpr bool kJoin = true;  // preserve_join_tree (main.cpp:21)

using Trip = std::array<int, 46>;

void by_saddle_id(std::vector<Trip>& t) {
  std::sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) { return a[45] < b[45]; });
}

struct Fields {
  int *ds, *as, *dds, *das;
  std::vector<std::pair<int, int>> mt, nt, dmt, dnt;
  std::vector<Branch> br, dbr;
};

// main.cpp:290-304 (and :342-355, :393-408): manifolds, CP lists and saddle
// maps of decp, triplets by saddle id.
void decp_cp(Fields& F, int nMax, int nMin) {
  ComputeDescendingManifold(decp_data, dec_vertex_type, F.dds);
  ComputeAscendingManifold(decp_data, dec_vertex_type, F.das);
  computeCP(F.dmt, F.dnt, dec_saddleTriplets, dec_saddle1Triplets, decp_data, F.dds, F.das,
            dec_vertex_type, dec_saddles2, dec_saddles1, dec_minimum, dec_maximum, dec_globalMin,
            0);
  computelargestSaddlesForMax(nMax, dec_saddleTriplets, dec_largestSaddlesForMax, &dec_maximum);
  computesmallestSaddlesForMin(nMin, dec_saddle1Triplets, dec_smallestSaddlesForMin,
                               &dec_minimum);
  by_saddle_id(dec_saddleTriplets);
  by_saddle_id(dec_saddle1Triplets);
}

// main.cpp:309-323: the wrong neighbours of every original saddle, in
// saddle-id order.
void wrong_neighbour_pass() {
  number_of_false_cases = 0;
  number_of_false_cases1 = 0;
  std::fill(wrong_neighbors, wrong_neighbors + num_Elements, 0);
  std::fill(wrong_neighbors_ds, wrong_neighbors_ds + num_Elements, 0);
  for (const Trip& t : saddleTriplets)
    get_wrong_split_neighbors(t[45], number_of_false_cases, decp_data);
  if (kJoin)
    for (const Trip& t : saddle1Triplets)
      get_wrong_join_neighbors(t[45], number_of_false_cases1, decp_data);
}

// main.cpp:326-338: the rank and saddle fixes of each side whose wrong
// neighbours are all fixed.
void rank_fixes() {
  if (number_of_false_cases == 0) {
    s_loops(kSplit);
    if (wrong_max_counter == 0) saddle_loops(kSplit);
  }
  if (number_of_false_cases1 == 0 && kJoin) {
    s_loops_join(kSplit);
    if (wrong_min_counter == 0) saddle_loops_join(kSplit);
  }
}

bool any_counter() {
  return number_of_false_cases > 0 || number_of_false_cases1 > 0 || count_f_max > 0 ||
         count_f_min > 0 || count_f_saddle > 0 || wrong_max_counter > 0 ||
         wrong_max_counter_2 > 0 || wrong_saddle_counter > 0 || wrong_min_counter > 0 ||
         wrong_min_counter_2 > 0 || wrong_saddle_counter_join > 0;
}

int* zeros(long long n) { return new int[n](); }

}  // namespace

int main(int argc, char** argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s field.f64 W H D er maxNeighbors out.bin\n", argv[0]);
    return 2;
  }
  const std::string path = argv[1];
  width = std::atoi(argv[2]);
  height = std::atoi(argv[3]);
  depth = std::atoi(argv[4]);
  const double er = std::atof(argv[5]);
  maxNeighbors = std::atoi(argv[6]);
  const std::string out = argv[7];
  num_Elements = width * height * depth;
  const long long N = num_Elements;

  // main.cpp:43-62: the caller's arrays
  input_data = new double[N];
  decp_data = new double[N];
  decp_data_copy = new double[N];
  d_deltaBuffer = new double[N];
  dec_vertex_type = new int[N];
  vertex_type = new int[N];
  all_max = new int[N];
  all_min = new int[N];
  all_saddle = new int[N];
  vertex_cells.resize(2 * std::max(0, width - 1) * std::max(0, height - 1));
  delta_counter.assign(N, 0);
  lowerStars = new int[(maxNeighbors + 1) * N];
  upperStars = new int[(maxNeighbors + 1) * N];
  dec_lowerStars = new int[(maxNeighbors + 1) * N];
  dec_upperStars = new int[(maxNeighbors + 1) * N];

  getdata(path, er, num_Elements, kSplit);

  // main.cpp:67-75: global extrema of the original, snapped in decp
  std::vector<int> order(N);
  getSortedIndexes(input_data, order.data());
  globalMin = order[0];
  const int globalMax = order[N - 1];
  decp_data[globalMin] = input_data[globalMin] - bound;
  decp_data[globalMax] = input_data[globalMax] - bound;

  adjacency = new int[N * maxNeighbors];
  computeAdjacency();
  get_vertex_traingle();

  Fields F;
  F.ds = new int[N];
  F.as = new int[N];
  classifyAllVertices(vertex_type, input_data, lowerStars, upperStars, 0);
  classifyAllVertices(dec_vertex_type, decp_data, dec_lowerStars, dec_upperStars, 1);
  ComputeDescendingManifold(input_data, vertex_type, F.ds);
  ComputeAscendingManifold(input_data, vertex_type, F.as);
  computeCP(F.mt, F.nt, saddleTriplets, saddle1Triplets, input_data, F.ds, F.as, vertex_type,
            saddles2, saddles1, minimum, maximum, globalMin, 0);
  const int nMax = (int)maximum.size(), nMin = (int)minimum.size();
  computelargestSaddlesForMax(nMax, saddleTriplets, largestSaddlesForMax, &maximum);
  computesmallestSaddlesForMin(nMin, saddle1Triplets, smallestSaddlesForMin, &minimum);
  by_saddle_id(saddleTriplets);
  by_saddle_id(saddle1Triplets);

  // main.cpp:122-131: every original saddle snapped in decp
  for (int s : saddles2) decp_data[s] = input_data[s] - bound;
  if (kJoin)
    for (int s : saddles1) decp_data[s] = input_data[s] - bound;

  c_loops(kSplit);

  // main.cpp:142-170
  or_saddle_max_map = new int[N * 4];
  or_saddle_min_map = new int[N * 4];
  wrong_neighbors = zeros(N);
  wrong_neighbors_index = zeros(N);
  wrong_neighbors_ds = zeros(N);
  wrong_neighbors_ds_index = zeros(N);
  for (int** a : {&wrong_rank_max, &wrong_rank_max_2, &wrong_rank_saddle, &wrong_rank_min,
                  &wrong_rank_min_2, &wrong_rank_saddle_join})
    *a = zeros(N);
  for (int** a : {&wrong_rank_max_index, &wrong_rank_max_index_2, &wrong_rank_saddle_index,
                  &wrong_rank_min_index, &wrong_rank_min_index_2, &wrong_rank_saddle_join_index})
    *a = zeros(2 * N);
  F.dds = new int[N];
  F.das = new int[N];

  ComputeDescendingManifold(decp_data, dec_vertex_type, F.dds);
  ComputeAscendingManifold(decp_data, dec_vertex_type, F.das);
  computeCP(F.dmt, F.dnt, dec_saddleTriplets, dec_saddle1Triplets, decp_data, F.dds, F.das,
            dec_vertex_type, dec_saddles2, dec_saddles1, dec_minimum, dec_maximum, dec_globalMin,
            0);
  by_saddle_id(dec_saddleTriplets);
  by_saddle_id(dec_saddle1Triplets);

  // main.cpp:195-207: original labels of every saddle's neighbours
#pragma omp parallel for
  for (long long i = 0; i < (long long)saddleTriplets.size(); i++)
    compute_Max_for_Saddle(saddleTriplets[i][45], input_data);
#pragma omp parallel for
  for (long long i = 0; i < (long long)saddle1Triplets.size(); i++)
    compute_Min_for_Saddle(saddle1Triplets[i][45], input_data);

  wrong_neighbour_pass();

  // main.cpp:230-274: the first rank fixes (the split side's CP state
  // recomputed without the ascending manifold, the join side's with both)
  if (number_of_false_cases == 0) {
    ComputeDescendingManifold(decp_data, dec_vertex_type, F.dds);
    computeCP(F.dmt, F.dnt, dec_saddleTriplets, dec_saddle1Triplets, decp_data, F.dds, F.das,
              dec_vertex_type, dec_saddles2, dec_saddles1, dec_minimum, dec_maximum,
              dec_globalMin, 0);
    computelargestSaddlesForMax(nMax, dec_saddleTriplets, dec_largestSaddlesForMax,
                                &dec_maximum);
    by_saddle_id(dec_saddleTriplets);
    s_loops(kSplit);
    if (wrong_max_counter == 0) saddle_loops(kSplit);
  }
  if (number_of_false_cases1 == 0 && kJoin) {
    ComputeDescendingManifold(decp_data, dec_vertex_type, F.dds);
    ComputeAscendingManifold(decp_data, dec_vertex_type, F.das);
    computeCP(F.dmt, F.dnt, dec_saddleTriplets, dec_saddle1Triplets, decp_data, F.dds, F.das,
              dec_vertex_type, dec_saddles2, dec_saddles1, dec_minimum, dec_maximum,
              dec_globalMin, 0);
    computesmallestSaddlesForMin(nMin, dec_saddle1Triplets, dec_smallestSaddlesForMin,
                                 &dec_minimum);
    by_saddle_id(dec_saddle1Triplets);
    s_loops_join(kSplit);
    if (wrong_min_counter == 0) saddle_loops_join(kSplit);
  }

  // main.cpp:277-446
  long long outer = 0;
  while (any_counter()) {
    if (++outer > 10000) {
      std::fprintf(stderr, "outer loop did not end\n");
      return 3;
    }
    r_loops(kSplit);
    for (int rep = 0; rep < 3; rep++) {
      c_loops(kSplit);
      decp_cp(F, nMax, nMin);
      wrong_neighbour_pass();
      if (rep < 2) {
        rank_fixes();
        continue;
      }
      // the last repetition only counts (main.cpp:432-444)
      if (number_of_false_cases == 0) {
        get_wrong_index_max();
        get_wrong_index_saddles();
      }
      if (number_of_false_cases1 == 0 && kJoin) {
        get_wrong_index_min();
        get_wrong_index_saddles_join();
      }
    }
  }

  // main.cpp:454-504: split trees of the corrected and the original field
  computeCP(F.dmt, F.dnt, dec_saddleTriplets, dec_saddle1Triplets, decp_data, F.dds, F.das,
            dec_vertex_type, dec_saddles2, dec_saddles1, dec_minimum, dec_maximum, dec_globalMin,
            1);
  compute_MergeT(F.dbr, dec_saddles2, dec_maximum, F.dmt, dec_saddleTriplets, decp_data,
                 num_Elements, dec_globalMin);
  computeCP(F.mt, F.nt, saddleTriplets, saddle1Triplets, decp_data, F.ds, F.as, vertex_type,
            saddles2, saddles1, minimum, maximum, globalMin, 1);
  compute_MergeT(F.br, saddles2, maximum, F.mt, saddleTriplets, input_data, num_Elements,
                 globalMin);
  long long changed = 0;
  for (long long i = 0; i < N; i++) changed += decp_data[i] != decp_data_copy[i];
  long long unmatched = 0;
  const size_t nb = std::min(F.br.size(), F.dbr.size());
  for (size_t b = 0; b < nb; b++) {
    const auto& v = F.br[b].vertices;
    const auto& w = F.dbr[b].vertices;
    for (size_t p = 0; p + 1 < v.size() && p + 1 < w.size(); p++)
      if (v[p].second != w[p].second || v[p + 1].second != w[p + 1].second) unmatched++;
  }
  saveArrayToBin(decp_data, N, out);
  std::printf("dropin: outer %lld changed %lld branches %zu/%zu unmatched %lld max %d min %d "
              "s2 %zu s1 %zu\n",
              outer, changed, F.br.size(), F.dbr.size(), unmatched, nMax, nMin, saddles2.size(),
              saddles1.size());
  return 0;
}
