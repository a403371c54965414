// extremz_cli -- command-line driver with the reference's argument form.
//
//   extremz_cli <path,W,H,D> <er> <compressor_id> [--decp decp.bin]
//               [--orig-f32] [--out fixed.bin] [--counts counts.bin]
//               [--cost prefix] [--cmp-bytes n] [--fixed-step q] [--neighbors 6|14]
//
// Same positional arguments as main.cpp:13-37: a raw float64 field (getdata,
// src/ExtreM.cpp:95-118; --orig-f32 for float32), the relative bound er
// (xi = (max - min) * er, src/ExtreM.cpp:136-138) and a compressor id that
// only names outputs.  The reference produces decp with SZ3 inside getdata
// (src/ExtreM.cpp:121-131); here decp is read from --decp, or, without it,
// made by the uniform quantiser decp = 2*xi*rint(orig/(2*xi)).  The corrected
// field is written as raw float64 to fixed.bin (main.cpp:507).  With --cost,
// the edit container of cost() (src/ExtreM.cpp:272-389) is written as
// <prefix>index.bin / <prefix>edits.bin and their zstd frames (.zst), and the
// statistics block of the reference's result file is appended to
// <prefix>stats.txt; --cmp-bytes is the lossy compressor's output size
// (SZ3's cmpSize in the reference).  --fixed-step q selects the GPU fork's
// fixed-step rule set (its edit_type 0 with q, src/ExtreM_cuda.cu:4958-4963).
// --neighbors 6 selects the CPU path's adjacency, maxNeighbors = 6 as main.cpp:20
// sets it (+-x and +-y on a field of depth 1); the default is the
// 14-neighbour stencil.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/extremz.h"

static bool read_raw(const std::string& path, void* dst, size_t bytes) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) {
    std::fprintf(stderr, "Cannot open file: %s\n", path.c_str());
    return false;
  }
  if ((size_t)f.tellg() != bytes) {
    std::fprintf(stderr, "File size does not match expected data size.\n");
    return false;
  }
  f.seekg(0);
  f.read(static_cast<char*>(dst), (std::streamsize)bytes);
  return (bool)f;
}

static bool write_raw(const std::string& path, const void* src, size_t bytes) {
  std::ofstream f(path, std::ios::binary);
  if (!f) {
    std::fprintf(stderr, "Error opening file for writing: %s\n", path.c_str());
    return false;
  }
  f.write(static_cast<const char*>(src), (std::streamsize)bytes);
  return (bool)f;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr,
                 "usage: %s <path,W,H,D> <er> <compressor_id> [--decp decp.bin] [--orig-f32] "
                 "[--out fixed.bin] [--counts counts.bin] [--cost prefix] [--cmp-bytes n] "
                 "[--fixed-step q] [--neighbors 6|14]\n",
                 argv[0]);
    return 2;
  }
  std::string dims = argv[1], path;
  std::istringstream iss(dims);
  int W = 0, H = 0, D = 0;
  char sep;
  if (!std::getline(iss, path, ',') || !(iss >> W >> sep) || sep != ',' || !(iss >> H >> sep) ||
      sep != ',' || !(iss >> D)) {
    std::fprintf(stderr, "Parsing error for dimensions\n");
    return 2;
  }
  double er = std::strtod(argv[2], nullptr);
  std::string compressor_id = argv[3];
  std::string decp_path, out_path = "fixed.bin", counts_path, cost_prefix;
  long long cmp_bytes = 0;
  int fixed_q = 0, adjacency = EXTREMZ_ADJ_STENCIL14;
  bool f32 = false;
  for (int i = 4; i < argc; i++) {
    std::string a = argv[i];
    if (a == "--decp" && i + 1 < argc) decp_path = argv[++i];
    else if (a == "--out" && i + 1 < argc) out_path = argv[++i];
    else if (a == "--counts" && i + 1 < argc) counts_path = argv[++i];
    else if (a == "--orig-f32") f32 = true;
    else if (a == "--cost" && i + 1 < argc) cost_prefix = argv[++i];
    else if (a == "--cmp-bytes" && i + 1 < argc) cmp_bytes = std::atoll(argv[++i]);
    else if (a == "--fixed-step" && i + 1 < argc) fixed_q = std::atoi(argv[++i]);
    else if (a == "--neighbors" && i + 1 < argc) {
      const int m = std::atoi(argv[++i]);
      if (m != 6 && m != 14) {
        std::fprintf(stderr, "--neighbors takes 6 or 14\n");
        return 2;
      }
      adjacency = m == 6 ? EXTREMZ_ADJ_PLANE4 : EXTREMZ_ADJ_STENCIL14;
    }
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return 2;
    }
  }
  const size_t N = (size_t)W * H * D;
  std::printf("Filepath: %s\nWidth: %d\nHeight: %d\nDepth: %d\nsize is:%zu\n", path.c_str(), W, H,
              D, N);
  std::vector<double> orig(N);
  if (f32) {
    std::vector<float> tmp(N);
    if (!read_raw(path, tmp.data(), N * 4)) return 1;
    for (size_t i = 0; i < N; i++) orig[i] = tmp[i];
  } else if (!read_raw(path, orig.data(), N * 8)) {
    return 1;
  }
  double mn = orig[0], mx = orig[0];
  for (double v : orig) {
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  const double xi = (mx - mn) * er;
  std::vector<double> decp(N);
  if (!decp_path.empty()) {
    if (!read_raw(decp_path, decp.data(), N * 8)) return 1;
  } else {
    const double q = 2.0 * xi;
    for (size_t i = 0; i < N; i++) {
      double d = q * std::rint(orig[i] / q);
      if (std::fabs(d - orig[i]) > xi) d -= (d > orig[i] ? q : -q);
      decp[i] = d;
    }
  }
  std::vector<double> before = decp;
  std::vector<uint8_t> counts(N);
  extremz_stats st;
  extremz_opts opts = {0, 0, 0, fixed_q, adjacency};
  int rc = extremz_correct(orig.data(), EXTREMZ_DTYPE_F64, decp.data(), W, H, D, xi, &opts, &st,
                           counts.data());
  if (rc != 0) {
    std::fprintf(stderr, "extremz_correct failed (%d): %s\n", rc, extremz_last_error());
    return 1;
  }
  size_t changed = 0;
  for (size_t i = 0; i < N; i++) changed += decp[i] != before[i];
  std::printf("bound: %.17g\nouter iterations: %lld\nc_loop iterations: %lld\n", xi,
              (long long)st.counts[EXTREMZ_STAT_OUTER_ITERS],
              (long long)st.counts[EXTREMZ_STAT_C_LOOP_ITERS]);
  std::printf("edit ratio: %.16f\nwall ms: %.3f\n", (double)changed / (double)N, st.wall_ms);
  if (!write_raw(out_path, decp.data(), N * 8)) return 1;
  if (!counts_path.empty() && !write_raw(counts_path, counts.data(), N)) return 1;
  if (!cost_prefix.empty()) {
    // raw streams (the two files of cost()) and their zstd frames
    std::vector<uint8_t> bitmap((3 * N + 7) / 8);
    std::vector<double> edits(N);
    int64_t n_edits = 0;
    rc = extremz_encode(orig.data(), before.data(), decp.data(), counts.data(), (int64_t)N, xi,
                        bitmap.data(), edits.data(), nullptr, nullptr, &n_edits);
    const int64_t icap = extremz_zstd_bound((int64_t)bitmap.size());
    const int64_t ecap = extremz_zstd_bound(8 * n_edits);
    if (rc != 0 || icap < 0 || ecap < 0) {
      std::fprintf(stderr, "cost failed (%d): %s\n", rc != 0 ? rc : (int)icap,
                   extremz_last_error());
      return 1;
    }
    std::vector<uint8_t> iz(icap), ez(ecap > 0 ? ecap : 1);
    extremz_cost_stats cs;
    const int64_t orig_bytes = (int64_t)N * (f32 ? 4 : 8);
    rc = extremz_cost(orig.data(), before.data(), decp.data(), counts.data(), (int64_t)N, xi,
                      orig_bytes, cmp_bytes, 3, iz.data(), icap, ez.data(), (int64_t)ez.size(),
                      &cs);
    if (rc != 0) {
      std::fprintf(stderr, "cost failed (%d): %s\n", rc, extremz_last_error());
      return 1;
    }
    if (!write_raw(cost_prefix + "index.bin", bitmap.data(), bitmap.size()) ||
        !write_raw(cost_prefix + "index.bin.zst", iz.data(), (size_t)cs.index_zst_bytes) ||
        !write_raw(cost_prefix + "edits.bin", edits.data(), (size_t)n_edits * 8) ||
        !write_raw(cost_prefix + "edits.bin.zst", ez.data(), (size_t)cs.edits_zst_bytes))
      return 1;
    std::ofstream sf(cost_prefix + "stats.txt", std::ios::app);
    sf << std::to_string(xi) << ":\n" << std::setprecision(17) << "related_error: " << er
       << "\nOCR: " << cs.ocr << "\nCR: " << cs.cr << "\nOBR: " << cs.obr << "\nBR: " << cs.br
       << "\npsnr: " << cs.psnr << "\nfixed_psnr: " << cs.fixed_psnr
       << "\nedit_ratio: " << cs.edit_ratio << "\nadditional_time: " << st.wall_ms / 1000.0
       << "\n\n";
    std::printf("%lld,%.17g\npsnr: %.17g\nfixed_psnr: %.17g\nindex zst bytes: %lld\n"
                "edits zst bytes: %lld\n",
                (long long)cs.changed, cs.edit_ratio, cs.psnr, cs.fixed_psnr,
                (long long)cs.index_zst_bytes, (long long)cs.edits_zst_bytes);
  }
  return 0;
}
