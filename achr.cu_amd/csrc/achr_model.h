// The model handle, the error convention and the warmup reader, shared by the
// host-only sources (achr_io.cpp, no HIP headers) and the HIP ones.
#pragma once
#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/achr.h"

struct achr_model {
  int n = 0, m = 0;
  std::vector<int> rowptr, colidx;
  std::vector<double> vals, lb, ub;
  std::vector<double> N;  // n x d column major
  std::vector<std::string> names;  // reaction (column) names of an MPS model
  int d = -1;             // -1: not computed yet
  // Columns n_file..n-1 of a model read with its rows honoured are auxiliary:
  // aux_row[k] is the row whose activity column n_file + k carries (its slack),
  // or -1 for the column fixed at 1 that carries the nonzero right-hand sides.
  int n_file = 0;
  std::vector<int> aux_row;
};

namespace achr_host {

// Records the text of a failure for achr_last_error() and returns code.
int fail(int code, const char* fmt, ...);

// Runs f() and turns a host allocation failure into ACHR_E_NOMEM, for entry
// points whose host buffers are sized by their inputs (files, models).
template <class F>
int no_throw(F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return fail(ACHR_E_NOMEM, "out of host memory");
  } catch (const std::length_error&) {
    return fail(ACHR_E_NOMEM, "host buffer too large");
  }
}

// Reads a warmup CSV (one line per reaction, comma-separated values, one
// column per point) into a malloc'ed n x W column-major array; n_expected 0
// accepts any line count, reported through n_out.
int read_warmup_csv(const char* path, int n_expected, double** warmup, int* W_out, int* n_out);

}  // namespace achr_host
