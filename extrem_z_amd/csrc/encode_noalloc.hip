// encode_noalloc.hip -- the edit encoding of cost() without any device
// allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// extremz_encode stages a flag and a position per voxel.  This variant keeps
// only one count per 1024-voxel block, in a file-scope device array
// (`scratch`), and recomputes each voxel's edit flag where it writes:
//   k_enc_count    edits per block          -> scratch[block]
//   k_enc_offsets  exclusive scan (1 block) -> scratch[block], total -> *n_edits
//   k_enc_write    block-local ballot scan of the flags, in voxel order
// plus k_bitmap3 for the 3-bit index bitmap.  The array holds
// kEncScratchInts ints (4 MiB + 256 B) on every device the library runs on,
// until the process ends; every call on a device shares it, so calls that
// overlap on one device race.  Nothing needs clearing between calls: every
// launch writes each entry it later reads.
#include <cmath>
#include <mutex>

#include "engine.h"

namespace ez {
namespace {

constexpr int kEncThreads = 256;
constexpr int kEncPer = 4;                             // voxels per thread
constexpr int kEncBlockVox = kEncThreads * kEncPer;    // voxels per block
constexpr long long kEncMaxN = 1LL << 30;              // EXTREMZ_ENCODE_NOALLOC_MAX_N
constexpr int kEncMaxBlocks = (int)(kEncMaxN / kEncBlockVox);
constexpr int kEncScratchInts = kEncMaxBlocks + 64;

static __device__ __align__(256) int scratch[kEncScratchInts];

// Device address of `scratch` per device, resolved once per device with that
// device current (each device holds its own copy of the array).
std::mutex g_scratch_mu;
int* g_scratch_ptr[64] = {};

int* scratch_on(int device) {
  if (device < 0 || device >= 64) throw HipError("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw HipError("hipGetSymbolAddress(scratch) returned no address");
    g_scratch_ptr[device] = static_cast<int*>(p);
  }
  return g_scratch_ptr[device];
}

// cost() writes an edit for a vertex that changed and was snapped (count 6),
// src/ExtreM.cpp:280-286.
__device__ __forceinline__ int enc_flag(const double* before, const double* after,
                                        const uint8_t* cnt, long long i) {
  return before[i] != after[i] && cnt[i] == 6;
}

__global__ void __launch_bounds__(kEncThreads) k_enc_count(const double* before,
                                                           const double* after,
                                                           const uint8_t* cnt, long long n,
                                                           int* blk) {
  __shared__ int wsum[kEncThreads / 64];
  const long long base = blockIdx.x * (long long)kEncBlockVox;
  int c = 0;
#pragma unroll
  for (int k = 0; k < kEncPer; k++) {
    const long long i = base + threadIdx.x + kEncThreads * k;
    if (i < n) c += enc_flag(before, after, cnt, i);
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0) wsum[threadIdx.x / 64] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kEncThreads / 64; w++) t += wsum[w];
    blk[blockIdx.x] = t;
  }
}

// Exclusive scan of the nblk block counts in place; the total goes to
// *n_edits.  One block of 1024 threads walks the counts in chunks.
__global__ void __launch_bounds__(1024) k_enc_offsets(int* blk, int nblk, int64_t* n_edits) {
  __shared__ int sh[1024];
  long long carry = 0;
  for (int base = 0; base < nblk; base += 1024) {
    const int i = base + (int)threadIdx.x;
    const int v = i < nblk ? blk[i] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nblk) blk[i] = (int)(carry + sh[threadIdx.x] - v);
    carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_edits = carry;
}

// Edits of the snapped vertices at their index-order positions: voxel
// base + t + 256 k (thread t, round k) follows every voxel of earlier rounds
// and the flagged voxels of threads below t in its round.
__global__ void __launch_bounds__(kEncThreads) k_enc_write(
    const double* orig, const double* before, const double* after, const uint8_t* cnt,
    long long n, double xi, const int* blk, double* edits, uint64_t* mant, int32_t* expo) {
  __shared__ int wsum[kEncThreads / 64];
  const long long base = blockIdx.x * (long long)kEncBlockVox;
  const int lane = __lane_id(), wave = threadIdx.x / 64;
  int off = blk[blockIdx.x];
  for (int k = 0; k < kEncPer; k++) {
    const long long i = base + threadIdx.x + kEncThreads * k;
    const int f = i < n ? enc_flag(before, after, cnt, i) : 0;
    const unsigned long long m = __ballot(f);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int before_wave = 0, round = 0;
    for (int w = 0; w < kEncThreads / 64; w++) {
      before_wave += w < wave ? wsum[w] : 0;
      round += wsum[w];
    }
    if (f) {
      const long long p = off + before_wave + below;
      // cost(): edits.push_back(-(decp_data_copy[i] - (input_data[i] - bound)))
      const double e = -(before[i] - (orig[i] - xi));
      const unsigned long long bits = (unsigned long long)__double_as_longlong(e);
      if (edits) edits[p] = e;
      // the split of src/ExtreM_cuda.cu:4466-4476
      if (mant) mant[p] = bits & 0xFFFFFFFFFFFFFull;
      if (expo) expo[p] = (int32_t)((bits >> 52) & 0x7FF) - 1023;
    }
    off += round;
    __syncthreads();  // wsum is rewritten by the next round
  }
}

}  // namespace
}  // namespace ez

extern "C" int extremz_encode_noalloc(extremz_ctx* ctx, const double* d_orig,
                                      const double* d_decp_before, const double* d_decp_after,
                                      const uint8_t* d_counts, int64_t n, double xi,
                                      uint8_t* d_index_bitmap, double* d_edits,
                                      uint64_t* d_mantissas, int32_t* d_exponents,
                                      int64_t* d_n_edits) {
  using namespace ez;
  if (!ctx || !d_orig || !d_decp_before || !d_decp_after || !d_counts || !d_index_bitmap ||
      !d_n_edits || n <= 0 || !std::isfinite(xi))
    return EXTREMZ_EINVAL;
  if (n > kEncMaxN) {
    g_err = "extremz_encode_noalloc: n exceeds EXTREMZ_ENCODE_NOALLOC_MAX_N (2^30 voxels), "
            "the capacity of its static scratch array";
    return EXTREMZ_EINVAL;
  }
  return guarded([&] {
    const int dev = ctx_device(ctx);
    hip_check(hipSetDevice(dev), "hipSetDevice");
    int* blk = scratch_on(dev);
    hipStream_t q = ctx_stream(ctx);
    const long long nbytes = (3 * n + 7) / 8;
    const int nblk = (int)((n + kEncBlockVox - 1) / kEncBlockVox);
    if (nblk > kEncMaxBlocks) throw HipError("block count exceeds the scratch array");
    k_bitmap3<<<2048, 256, 0, q>>>(d_decp_before, d_decp_after, d_counts, n, d_index_bitmap,
                                   nbytes);
    k_enc_count<<<nblk, kEncThreads, 0, q>>>(d_decp_before, d_decp_after, d_counts, n, blk);
    k_enc_offsets<<<1, 1024, 0, q>>>(blk, nblk, d_n_edits);
    if (d_edits || d_mantissas || d_exponents)
      k_enc_write<<<nblk, kEncThreads, 0, q>>>(d_orig, d_decp_before, d_decp_after, d_counts, n,
                                               xi, blk, d_edits, d_mantissas, d_exponents);
    hip_check(hipGetLastError(), "extremz_encode_noalloc launch");
    return EXTREMZ_OK;
  });
}
