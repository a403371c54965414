/* include/api/ExtreM.h -- the reference library's API (include/api/ExtreM.h:1-127
 * of the reference), served by libExtreMz over the MI355X engine.
 *
 * Link-compatible with the reference: the same globals and the same C-linkage
 * functions with the same C++ parameter types, so a caller written against
 * the reference (its main.cpp) compiles against this header and links against
 * libExtreMz unchanged.  The one change is the direction table: the reference
 * header declares `directions[78]` while its library defines 42 entries
 * (src/ExtreM.cpp:16), which no program can link; it is declared as the 42
 * entries the library defines.
 *
 * Semantics (each function cites the reference definition it replaces in
 * extrem_z_amd/csrc/compat.hip):
 *  - The caller allocates the arrays and sets width / height / depth /
 *    maxNeighbors / num_Elements, as with the reference.
 *  - The fields and the correction live on the GPU.  After every call the
 *    arrays the function writes in the reference hold its results (decp_data,
 *    delta_counter, the counters, the index lists, the manifold and star
 *    arrays, the CP vectors), except the per-vertex scratch the correction
 *    steps keep on the device (d_deltaBuffer, all_max / all_min / all_saddle,
 *    the wrong_rank_* index pairs and flags).
 *  - decp_data may be edited by the caller between calls (main.cpp's
 *    pre-snap does so); input_data is read-only once the first correction
 *    step has run.
 *  - maxNeighbors == 6 on a field of depth 1 is the reference CPU path's
 *    plane adjacency (+-x, +-y); any other field uses the 14-direction
 *    stencil of the reference's GPU path.  The arrays the caller sizes by
 *    maxNeighbors (adjacency, lowerStars / upperStars) hold the first
 *    maxNeighbors directions of the table, as the reference's fill loops do.
 *  - tree_type must be "split" (main.cpp:63).
 *  - Errors are printed to std::cerr, as the reference does; a missing GPU
 *    throws std::runtime_error.
 */
#ifndef EXTREM_H
#define EXTREM_H
#include <cstdint>
#include <stddef.h>
#include <string>
#include <array>
#include <vector>
#include <unordered_map>
#include <atomic>

/* One branch of the split tree: (order, vertex id) pairs, the first one the
 * branch's maximum; parentBranch points into the same branch vector. */
struct Branch {
    std::vector<std::pair<int, int>> vertices;
    Branch *parentBranch = nullptr;

    void print() const;
};

extern "C" {
    /* The 14 stencil directions (x, y, z) and the 26-neighbour offset table. */
    extern int directions[42], direction_to_index_mapping[26][3];

    /* Shape, adjacency width and scalar state. */
    extern double* d_deltaBuffer;
    extern int width, height, depth, maxNeighbors, num_Elements, number_of_false_cases;
    extern std::atomic_int count_f_max, count_f_min, count_f_saddle;
    extern int wrong_max_counter, globalMin, dec_globalMin, wrong_max_counter_2, wrong_saddle_counter;
    extern int wrong_saddle_counter_join;

    /* Per-vertex arrays (caller-allocated). */
    extern int* adjacency, *dec_vertex_type, *vertex_type, *all_max, *all_min, *all_saddle;
    extern double *decp_data, *input_data, *decp_data_copy;
    extern double bound;
    extern int *or_saddle_max_map, *wrong_neighbors, *wrong_neighbors1, *wrong_neighbors_index, *wrong_rank_max, *wrong_rank_max_index, *wrong_rank_saddle, *wrong_rank_saddle_index, *wrong_rank_max_2, *wrong_rank_max_index_2;
    extern int *wrong_rank_ds_saddle, *wrong_rank_ds_saddle_index;
    extern int *or_saddle_min_map, *wrong_neighbors_ds, *wrong_neighbors_ds_index, *wrong_rank_min, *wrong_rank_min_index, *wrong_rank_min_2, *wrong_rank_min_index_2;
    extern int number_of_false_cases1, wrong_min_counter, wrong_min_counter_2;

    /* Critical-point structures (library-owned containers). */
    extern std::vector<std::array<int, 46>> saddleTriplets, dec_saddleTriplets;
    extern std::vector<std::array<int, 46>> saddle1Triplets, dec_saddle1Triplets;
    extern std::vector<std::vector<int>> vertex_cells;
    extern std::vector<int> saddles2, maximum, dec_saddles2, dec_maximum, delta_counter;
    extern std::vector<int> saddles1, minimum, dec_saddles1, dec_minimum;
    extern std::unordered_map<int, int> largestSaddlesForMax, dec_largestSaddlesForMax;
    extern std::unordered_map<int, int> smallestSaddlesForMin, dec_smallestSaddlesForMin;
    extern int *lowerStars, *upperStars, *dec_lowerStars, *dec_upperStars;

    extern int *wrong_rank_saddle_join, *wrong_rank_saddle_join_index;

    /* I/O and set-up. */
    void saveArrayToBin(const double* arr, size_t size, const std::string& filename);
    void get_vertex_traingle();
    void getdata(const std::string &filename,
                const double er,
                int data_size, std::string type);

    /* Phases (a) and (b): classification and manifolds. */
    int ComputeDescendingManifold(const double *offset, const int *vertex_type, int *DS_M);
    int ComputeAscendingManifold(const double *offset, const int *vertex_type, int *AS_M);
    void classifyAllVertices(int *vertex_type, const double* heightMap, int *lowerStars, int *upperStars, int type);
    void computeAdjacency();
    void c_loops(std::string type);
    int classifyVertex(const int vertexId, const double *offset,
                                        const int *desManifold, const int *ascManifold,
                                        int *types);

    int computeNumberOfLinkComponents(const int* linkVertices, const int nLinkVertices);
    void sortAndRemoveDuplicates(std::array<int, 46> &triplet);
    void computeCP(std::vector<std::pair<int, int>> &maximaTriplets, std::vector<std::pair<int, int>> &minimumTriplets,
                    std::vector<std::array<int, 46>> &saddleTriplets, std::vector<std::array<int, 46>> &saddle1Triplets,
                    const double *offset, const int *desManifold, const int *ascManifold,
                    const int *vertex_type_tmp, std::vector<int> &saddles2,
                    std::vector<int> &saddles1, std::vector<int> &minimum,
                    std::vector<int> &maximum, int &globalMin, int translation = 0);

    void getSortedIndexes(const double* vec, int* indexes);
    void findAscPaths(  std::vector<std::array<int, 46>> &saddleTriplets,
                        std::vector<int> &maximaLocalToGlobal,
                        std::vector<int> &saddlesLocalToGlobal,
                        const int *saddles2, const int *maximum,
                        const double *offset,
                        const int *desManifold, const int *ascManifold, int &counter_tmp);

    int constructPersistencePairs(
                        std::vector<std::pair<int, int>> &pairs,
                        std::vector<std::pair<int, int>> &maximaTriplets,
                        std::vector<std::array<int, 46>> &saddleTriplets,
                        const int* maximum,
                        const int* saddles2,
                        const int nMax,
                        int globalMin);

    /* Phase (c): the correction steps main.cpp:139-446 drives. */
    void compute_Max_for_Saddle(int saddle, const double *offset);
    void compute_Min_for_Saddle(int saddle, const double *offset);
    void get_wrong_split_neighbors(int saddle, int &num_false_cases, const double *offset);
    void get_wrong_join_neighbors(int saddle, int &num_false_cases, const double *offset);
    int fixpath(int i, int direction, std::string type);
    void get_false_criticle_points();
    void r_loops(std::string type);
    void s_loops(std::string type);
    void s_loops_join(std::string type);

    void get_wrong_index_max();
    void compute_MergeT(std::vector<Branch> &branches,
                        const std::vector<int> saddles2,
                        const std::vector<int> maximum,
                        std::vector<std::pair<int, int>> &maximaTriplets,
                        std::vector<std::array<int, 46>> &saddleTriplets,
                        const double *offset,
                        const int data_size, const int globalMin);
    void computelargestSaddlesForMax(const int nMax,
                                    const std::vector<std::array<int, 46>> &saddleTriplets,
                                    std::unordered_map<int, int> &largestSaddlesForMax,
                                    const std::vector<int> *saddles2);
    void computesmallestSaddlesForMin(const int nMin,
                                    const std::vector<std::array<int, 46>> &saddleTriplets,
                                    std::unordered_map<int, int> &smallestSaddlesForMin,
                                    const std::vector<int> *minimum);
    void reorderVectorByTemparray(std::vector<std::pair<int, int>>& vec, const std::vector<int>& temparray);
    void get_wrong_index_saddles();
    void get_wrong_index_saddles_join();
    void saddle_loops(std::string type);
    void saddle_loops_join(std::string type);
    void get_wrong_index_min();
    void mappath(int* DS_M, int* AS_M, const double* offset);
    int computePathCompressionSingle(
        int *const segmentation,
        const bool computeAscending,
        const double *const offset) ;

}

#endif /* EXTREM_H */
