// ACHRCuda — drop-in command line of the MI355X sampler.
//
//   ACHRCuda <model.mps> <warmup.csv> <nFiles> <pointsPerFile> <stepsPerPoint>
//
// Same arguments and outputs as the reference's main (ACHRcu/ACHRCuda.cu:504-802):
// reads the MPS model and the warmup CSV, computes the null space of S and
// the warmup centre, then for each file k runs pointsPerFile chains from the
// centre for stepsPerPoint accepted steps and writes File<k> (one line per
// reaction, comma-separated "%f" values, one column per chain) to the working
// directory. Unlike the reference it checks its arguments and exits non-zero
// on any failure.
//
// Options beyond the reference, all through the environment so the positional
// interface stays unchanged:
//   ACHR_SEED=<u64>        Philox key (default 0: runs are reproducible)
//   ACHR_LANES=<16..1024>  threads per chain (default: autotuned schedule)
//   ACHR_REPROJECT=1|2     1 = per chain in the step kernel, 2 = batched GEMM
//   ACHR_NULLSPACE=<file>  read N (n x d, column major, raw little-endian
//                          doubles) instead of computing it
//   ACHR_MPS_ROWS=1        honour the file's row types, right-hand sides and
//                          ranges (the set createWarmupPts solves over) instead
//                          of sampling S x = 0 over every row as the reference
//                          does; File<k> still holds the file's columns only
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "achr.h"

static int die(const char* what) {
  std::fprintf(stderr, "ACHRCuda: %s: %s\n", what, achr_last_error());
  return 1;
}

int main(int argc, char** argv) {
  // The reference's five arguments; an optional sixth, --lossless, also
  // writes each file's points bit-exact to File<k>.bin.
  const bool lossless = argc == 7 && std::strcmp(argv[6], "--lossless") == 0;
  if (argc != 6 && !lossless) {
    std::fprintf(stderr, "Usage: %s <model.mps> <warmup.csv> <nFiles> <pointsPerFile> <stepsPerPoint> [--lossless]\n",
                 argv[0]);
    return 2;
  }
  const auto t0 = std::chrono::steady_clock::now();
  char* end = nullptr;
  const long nFiles = std::strtol(argv[3], &end, 10);
  if (*end || nFiles < 0) return std::fprintf(stderr, "ACHRCuda: bad nFiles '%s'\n", argv[3]), 2;
  const long P = std::strtol(argv[4], &end, 10);
  if (*end || P <= 0 || P > (1L << 30)) return std::fprintf(stderr, "ACHRCuda: bad pointsPerFile '%s'\n", argv[4]), 2;
  const long S = std::strtol(argv[5], &end, 10);
  if (*end || S < 0 || S > (1L << 30)) return std::fprintf(stderr, "ACHRCuda: bad stepsPerPoint '%s'\n", argv[5]), 2;
  std::printf("Total count is %lld \n", static_cast<long long>(nFiles) * P * S);
  std::printf("\nThe model supplied is %s\n", argv[1]);

  achr_model* model = nullptr;
  const char* rows_env = std::getenv("ACHR_MPS_ROWS");
  const bool rows = rows_env && std::strcmp(rows_env, "1") == 0;
  if ((rows ? achr_model_read_mps_rows(argv[1], &model) : achr_model_read_mps(argv[1], &model)) != ACHR_OK)
    return die("reading the MPS model");
  int n = 0, m = 0, nnz = 0, d = 0, nf = 0;
  achr_model_dims(model, &n, &m, &nnz, nullptr);
  achr_model_file_columns(model, &nf);
  if (nf < n) std::printf("%d slack / right-hand-side columns carry the rows' types and right-hand sides\n", n - nf);
  std::printf("nRxns %d\nnMets %d\nnnz is %d \n", n, m, nnz);

  if (const char* npath = std::getenv("ACHR_NULLSPACE")) {
    FILE* f = std::fopen(npath, "rb");
    if (!f) return std::fprintf(stderr, "ACHRCuda: cannot open %s\n", npath), 1;
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (bytes % (8L * n)) return std::fprintf(stderr, "ACHRCuda: %s is not n x d doubles\n", npath), 1;
    d = static_cast<int>(bytes / (8L * n));
    std::vector<double> N(static_cast<size_t>(n) * d);
    if (std::fread(N.data(), 8, N.size(), f) != N.size()) return std::fprintf(stderr, "ACHRCuda: short read\n"), 1;
    std::fclose(f);
    if (achr_model_set_nullspace(model, N.data(), d) != ACHR_OK) return die("setting the null space");
  } else if (achr_model_compute_nullspace(model, &d) != ACHR_OK) {
    return die("computing the null space");
  }

  achr_options opt;
  achr_options_default(&opt);
  if (const char* v = std::getenv("ACHR_LANES")) opt.lanes = std::atoi(v);
  if (const char* v = std::getenv("ACHR_REPROJECT")) opt.reproject_mode = std::atoi(v);
  uint64_t seed = 0;
  if (const char* v = std::getenv("ACHR_SEED")) seed = std::strtoull(v, nullptr, 10);

  achr_sampler* sampler = nullptr;
  if (achr_sampler_create_csv(model, argv[2], &opt, &sampler) != ACHR_OK) return die("reading the warmup points");
  int lanes = 0, mode = 0;
  achr_sampler_schedule(sampler, &lanes, &mode);
  const double t_setup =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("null space dimension %d, %d threads per chain, %s reprojection\n", d, lanes,
              mode == 1 ? "in-kernel" : "batched");
  std::printf("Null space done in %.5f seconds.\n", t_setup);

  std::vector<double> pts(static_cast<size_t>(n) * P);
  for (long k = 0; k < nFiles; ++k) {
    std::printf("File %ld\n", k);
    if (S == 0) {
      // The reference clears its points buffer before every file
      // (ACHRCuda.cu:739) and writes a chain's point only inside the step
      // loop (:322, :350), so a file of zero steps holds zeros. (The library's
      // achr_sample returns each chain's start point instead.)
      std::fill(pts.begin(), pts.end(), 0.0);
    } else {
      if (achr_sample(sampler, static_cast<int>(P), static_cast<int>(S), seed, static_cast<uint32_t>(k), 0) != ACHR_OK)
        return die("sampling");
      if (achr_sample_points(sampler, pts.data()) != ACHR_OK) return die("fetching points");
    }
    const std::string name = "File" + std::to_string(k);
    // reaction major: the file's columns are the first nf rows
    if (achr_write_points_csv(name.c_str(), pts.data(), nf, static_cast<int>(P)) != ACHR_OK)
      return die("writing points");
    if (lossless && achr_write_points_bin((name + ".bin").c_str(), pts.data(), nf, static_cast<int>(P)) != ACHR_OK)
      return die("writing lossless points");
  }
  const double t_all = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Sampling done in %.5f seconds.\n", t_all);
  achr_sampler_destroy(sampler);
  achr_model_destroy(model);
  return 0;
}
