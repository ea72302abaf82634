// Exact finish of warmup LPs left unconverged by the batched first-order
// solver (achr_polish.cpp). Host-only.
#pragma once
#include <functional>

namespace achr_host {

// min c^T x s.t. x = N z (S x = 0), lb <= x <= ub, by a primal active-set
// method in z. Nr: the orthonormal null-space basis, row major (n x d, row i
// at Nr + i*d). x: in, the start (pulled into the bounds toward xfeas, a
// point within the bounds and on S x = 0 to rounding); out, the solution.
// Returns 0 at a KKT point, 1 at the iteration cap, 2 without a feasible
// start, 3 for an unbounded direction.
int polish_lp(int n, int d, const double* Nr, const double* lb, const double* ub, const double* c,
              const double* xfeas, double* x, int max_iters);

// Runs job(0 .. count-1) on up to `threads` host threads.
void polish_many(int count, int threads, const std::function<void(int)>& job);

}  // namespace achr_host
