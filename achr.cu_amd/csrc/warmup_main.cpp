// createWarmupPts — the warmup generator's command line
// (VFWarmup/createWarmupPtsGLPK.c:185-310), with the LPs solved on the GPU.
//
//   createWarmupPts <model.mps> <nPts> [-1]
//
// Writes <model><nPts>warmup.csv (one line per reaction, one "%f" column per
// point: the file ACHRCuda reads) and <model>_fva_bounds.csv ("reaction,min,
// max") next to the model, as the reference does (<model> is the path without
// its ".mps"). nPts below 2 n is raised to 2 n and an odd nPts to the next
// even number (a minimum and a maximum per objective). The third argument -1
// (the reference's GLPK scaling switch) is accepted; the GPU solver always
// rescales. The LPs honour the file's row types, right-hand sides and ranges
// as GLPK does for the reference (an E / L / G row is S_i x = / <= / >= b_i);
// for the usual S x = 0 model that is the sampler's own set. Environment:
// ACHR_SEED (random objectives), ACHR_WARMUP_ITERS (iteration cap, default
// 40,000), ACHR_WARMUP_TOL (default 1e-9), ACHR_WARMUP_POLISH (0 leaves LPs
// unconverged at the cap as they stopped; default 1 finishes them exactly).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "achr.h"

static int die(const char* what) {
  std::fprintf(stderr, "createWarmupPts: %s: %s\n", what, achr_last_error());
  return 1;
}

int main(int argc, char** argv) {
  if (argc < 3 || argc > 4) {
    std::fprintf(stderr, "Usage: %s <datafile.mps> <nPts> [-1]\n", argv[0]);
    return 2;
  }
  char* end = nullptr;
  long nPts = std::strtol(argv[2], &end, 10);
  if (*end || nPts <= 0 || nPts > (1L << 30)) return std::fprintf(stderr, "createWarmupPts: bad nPts '%s'\n", argv[2]), 2;
  if (argc == 4 && std::strcmp(argv[3], "-1") != 0)
    return std::fprintf(stderr, "createWarmupPts: third argument must be -1\n"), 2;
  std::string base = argv[1];
  if (base.size() >= 4 && base.compare(base.size() - 4, 4, ".mps") == 0) base.resize(base.size() - 4);
  std::printf("\nThe model supplied is %s\n", argv[1]);
  const auto t0 = std::chrono::steady_clock::now();
  achr_model* model = nullptr;
  if (achr_model_read_mps_rows(argv[1], &model) != ACHR_OK) return die("reading the MPS model");
  int n = 0, m = 0, nnz = 0, d = 0, nf = 0;
  achr_model_dims(model, &n, &m, &nnz, &d);
  achr_model_file_columns(model, &nf);
  if (nf < n) std::printf("%d slack / right-hand-side columns carry the rows' types and right-hand sides\n", n - nf);
  if (nPts < 2L * nf) {
    std::printf("Warning: nPts=%ld is less than 2*nRxns=%d, setting nPts=%d\n", nPts, 2 * nf, 2 * nf);
    nPts = 2L * nf;
  }
  if (nPts % 2) ++nPts;
  std::printf("Creating %ld warmup points! \n", nPts);
  achr_warmup_options opt;
  achr_warmup_options_default(&opt);
  if (const char* s = std::getenv("ACHR_SEED")) opt.seed = std::strtoull(s, nullptr, 10);
  if (const char* s = std::getenv("ACHR_WARMUP_ITERS")) opt.max_iters = std::atoi(s);
  if (const char* s = std::getenv("ACHR_WARMUP_TOL")) opt.tol = std::atof(s);
  if (const char* s = std::getenv("ACHR_WARMUP_POLISH")) opt.polish = std::atoi(s);
  std::vector<double> W(static_cast<size_t>(n) * nPts), fva(2 * static_cast<size_t>(n)), rep(6);
  std::vector<int> conv(nPts);
  if (achr_warmup_generate(model, static_cast<int>(nPts), &opt, W.data(), fva.data(), conv.data(), rep.data()) !=
      ACHR_OK)
    return die("solving the LPs");
  std::printf("%.0f of %ld LPs converged in %.0f iterations, %.3f s", rep[1], nPts, rep[0], rep[2] * 1e-3);
  if (rep[4] > 0) std::printf(" (%.0f finished by the exact polish, %.3f s)", rep[4], rep[5] * 1e-3);
  std::printf("\n");
  if (rep[1] < nPts)
    std::printf("Warning: %ld LPs stopped at the iteration cap; their points are approximate "
                "(raise ACHR_WARMUP_ITERS)\n", nPts - static_cast<long>(rep[1]));
  const std::string fva_name = base + "_fva_bounds.csv";
  if (achr_write_fva_csv(fva_name.c_str(), model, fva.data()) != ACHR_OK) return die("writing the FVA bounds");
  std::printf("Raw FVA bounds written to %s\n", fva_name.c_str());
  // one line per reaction of the file, one column per point
  std::vector<double> rows(static_cast<size_t>(nf) * nPts);
  for (long b = 0; b < nPts; ++b)
    for (int i = 0; i < nf; ++i) rows[static_cast<size_t>(i) * nPts + b] = W[static_cast<size_t>(b) * n + i];
  const std::string name = base + std::to_string(nPts) + "warmup.csv";
  if (achr_write_points_csv(name.c_str(), rows.data(), nf, static_cast<int>(nPts)) != ACHR_OK)
    return die("writing the warmup points");
  const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Warmup points written to %s in %.3f s\n", name.c_str(), t);
  achr_model_destroy(model);
  return 0;
}
