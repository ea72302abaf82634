// Exact finish for warmup LPs that the batched first-order solver leaves
// unconverged (achr_warmup.hip). The reference solves every LP with GLPK's
// simplex (VFWarmup/createWarmupPtsGLPK.c:100-175); a first-order method can
// stall where an optimum is tiny against bounds of +-1000. Here each such LP
//   min c^T x  s.t.  S x = 0,  lb <= x <= ub
// is solved in the coordinates of an orthonormal null-space basis N (x = N z,
// so S x = 0 holds throughout) by a primal active-set method: from a feasible
// start (the solver's point, pulled into the bounds toward a feasible point),
// move along the projected steepest descent direction -(I - Q Q^T) N^T c to
// the first blocking bound, which joins the active set (Q R: its rows, by
// Gram-Schmidt with reorthogonalisation); on a face where the projected
// gradient vanishes, the multipliers of the active bounds decide: all of the
// right sign is optimal (KKT), otherwise the bound with the most negative one
// leaves (the smallest index after a zero step, Bland's rule against
// cycling). The final point is moved onto its active bounds exactly. Host
// code with no device work (the LPs are small and few), threads over LPs.
#include "achr_polish.h"

#include <algorithm>
#include <new>
#include <cmath>
#include <thread>
#include <vector>

namespace achr_host {

namespace {

struct Active {
  int d;
  std::vector<int> idx, side;    // bound index, +1 upper / -1 lower / 0 both (fixed)
  std::vector<double> Q, R;      // Q: d x k (column major), R: upper k x k stored with leading dimension d
  int k() const { return static_cast<int>(idx.size()); }
};

double dot(const double* a, const double* b, int d) {
  double s = 0.0;
  for (int j = 0; j < d; ++j) s += a[j] * b[j];
  return s;
}

// Adds row a (length d) to the orthonormal basis; false if its part outside
// the span of the active rows is below rel |a| or 1e-11 (then nothing
// changes).
bool add_row(Active& A, const double* a, int i, int side, double rel = 1e-8) {
  const int d = A.d, k = A.k();
  if (k >= d) return false;
  std::vector<double> q(a, a + d), r(k + 1, 0.0);
  const double an = std::sqrt(dot(a, a, d));
  if (!(an > 0.0)) return false;
  for (int pass = 0; pass < 2; ++pass)
    for (int l = 0; l < k; ++l) {
      const double* ql = A.Q.data() + static_cast<size_t>(l) * d;
      const double c = dot(ql, q.data(), d);
      r[l] += c;
      for (int j = 0; j < d; ++j) q[j] -= c * ql[j];
    }
  const double rho = std::sqrt(dot(q.data(), q.data(), d));
  // also an absolute floor: rows of N are at most unit length, and a tiny
  // pivot of R would make the multipliers meaningless
  if (!(rho > rel * an) || !(rho > 1e-11)) return false;
  for (double& v : q) v /= rho;
  r[k] = rho;
  A.Q.insert(A.Q.end(), q.begin(), q.end());
  for (int l = 0; l <= k; ++l) A.R[static_cast<size_t>(k) * d + l] = r[l];
  A.idx.push_back(i);
  A.side.push_back(side);
  return true;
}

// Removes active row l: its column leaves R, which Givens rotations of
// adjacent rows (applied to the columns of Q as well) return to triangular
// form; the last column of Q then drops out.
void drop_row(Active& A, int l) {
  const int d = A.d, k = A.k();
  double* R = A.R.data();
  for (int c = l; c < k - 1; ++c)
    for (int r = 0; r <= c + 1; ++r) R[static_cast<size_t>(c) * d + r] = R[static_cast<size_t>(c + 1) * d + r];
  for (int r = 0; r < d; ++r) R[static_cast<size_t>(k - 1) * d + r] = 0.0;
  for (int j = l; j < k - 1; ++j) {
    const double a = R[static_cast<size_t>(j) * d + j], b = R[static_cast<size_t>(j) * d + j + 1];
    const double h = std::hypot(a, b);
    if (h == 0.0) continue;
    const double cs = a / h, sn = b / h;
    for (int c = j; c < k - 1; ++c) {
      double& u = R[static_cast<size_t>(c) * d + j];
      double& v = R[static_cast<size_t>(c) * d + j + 1];
      const double nu = cs * u + sn * v, nv = -sn * u + cs * v;
      u = nu;
      v = nv;
    }
    R[static_cast<size_t>(j) * d + j + 1] = 0.0;
    double* qj = A.Q.data() + static_cast<size_t>(j) * d;
    double* qk = A.Q.data() + static_cast<size_t>(j + 1) * d;
    for (int r = 0; r < d; ++r) {
      const double u = qj[r], v = qk[r];
      qj[r] = cs * u + sn * v;
      qk[r] = -sn * u + cs * v;
    }
  }
  A.Q.resize(static_cast<size_t>(k - 1) * d);
  A.idx.erase(A.idx.begin() + l);
  A.side.erase(A.side.begin() + l);
}

}  // namespace

int polish_lp(int n, int d, const double* Nr, const double* lb, const double* ub, const double* c,
              const double* xfeas, double* x, int max_iters) {
  if (d == 0) {
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    return 0;
  }
  // feasibility tolerance of each bound, relative to the bound's own size
  auto tol_lo = [&](int i) { return 1e-9 * std::max(1.0, std::fabs(lb[i])); };
  auto tol_hi = [&](int i) { return 1e-9 * std::max(1.0, std::fabs(ub[i])); };
  // z of the start and of the feasible point (N has orthonormal columns)
  std::vector<double> z(d, 0.0), zf(d, 0.0);
  for (int i = 0; i < n; ++i) {
    const double* a = Nr + static_cast<size_t>(i) * d;
    for (int j = 0; j < d; ++j) {
      z[j] += a[j] * x[i];
      zf[j] += a[j] * xfeas[i];
    }
  }
  auto eval = [&](const std::vector<double>& zz, std::vector<double>& xx) {
    for (int i = 0; i < n; ++i) xx[i] = dot(Nr + static_cast<size_t>(i) * d, zz.data(), d);
  };
  std::vector<double> xv(n), xf(n);
  eval(z, xv);
  eval(zf, xf);
  // into the bounds: the smallest move toward the feasible point that clears
  // every violation beyond tolerance
  double alpha = 0.0;
  for (int i = 0; i < n; ++i) {
    double v = 0.0, room = 0.0;
    if (xv[i] > ub[i] + tol_hi(i)) {
      v = xv[i] - ub[i];
      room = ub[i] - xf[i];
    } else if (xv[i] < lb[i] - tol_lo(i)) {
      v = lb[i] - xv[i];
      room = xf[i] - lb[i];
    } else {
      continue;
    }
    alpha = std::max(alpha, room > 0.0 ? v / (v + room) : 1.0);
  }
  if (alpha > 0.0) {
    alpha = std::min(1.0, alpha * (1.0 + 1e-12));
    for (int j = 0; j < d; ++j) z[j] += alpha * (zf[j] - z[j]);
    eval(z, xv);
  }
  for (int i = 0; i < n; ++i)
    if (xv[i] > ub[i] + 10 * tol_hi(i) || xv[i] < lb[i] - 10 * tol_lo(i)) return 2;  // no feasible start
  // objective in z
  std::vector<double> g(d, 0.0);
  for (int i = 0; i < n; ++i)
    if (c[i] != 0.0) {
      const double* a = Nr + static_cast<size_t>(i) * d;
      for (int j = 0; j < d; ++j) g[j] += c[i] * a[j];
    }
  const double gn = std::sqrt(dot(g.data(), g.data(), d));
  Active A;
  A.d = d;
  A.R.assign(static_cast<size_t>(d) * d, 0.0);
  // 0 inactive, 1 active, 2 nearly in the span of the active rows (still a
  // blocking candidate, added with a looser test when it blocks), 3 in it
  std::vector<char> state(n, 0);
  for (int i = 0; i < n; ++i)
    if (lb[i] == ub[i] && add_row(A, Nr + static_cast<size_t>(i) * d, i, 0)) state[i] = 1;
  for (int i = 0; i < n; ++i) {
    if (state[i] || lb[i] == ub[i]) continue;
    const int s = xv[i] >= ub[i] - tol_hi(i) ? 1 : (xv[i] <= lb[i] + tol_lo(i) ? -1 : 0);
    if (s != 0 && add_row(A, Nr + static_cast<size_t>(i) * d, i, s)) state[i] = 1;
  }
  std::vector<double> p(d), Np(n), w(d);
  bool zero_step = false;
  int it = 0;
  for (; it < max_iters; ++it) {
    const int k = A.k();
    // p = -(g - Q Q^T g)
    for (int j = 0; j < d; ++j) p[j] = -g[j];
    for (int pass = 0; pass < 2; ++pass)
      for (int l = 0; l < k; ++l) {
        const double* ql = A.Q.data() + static_cast<size_t>(l) * d;
        const double cq = dot(ql, p.data(), d);
        for (int j = 0; j < d; ++j) p[j] -= cq * ql[j];
      }
    const double pn = std::sqrt(dot(p.data(), p.data(), d));
    if (pn > 1e-12 * gn) {
      for (int i = 0; i < n; ++i) Np[i] = dot(Nr + static_cast<size_t>(i) * d, p.data(), d);
      // a bound takes part unless its row is orthogonal to p to rounding
      // (rows of N have norm <= 1)
      const double thr = 1e-13 * pn;
      int jb = -1;
      double tb = INFINITY;
      for (int i = 0; i < n; ++i) {
        if ((state[i] != 0 && state[i] != 2) || !(std::fabs(Np[i]) > thr)) continue;
        double t = Np[i] > 0.0 ? (ub[i] - xv[i]) / Np[i] : (lb[i] - xv[i]) / Np[i];
        if (t < 0.0) t = 0.0;
        if (t < tb) {
          tb = t;
          jb = i;
        }
      }
      if (jb < 0) return 3;  // unbounded direction: not a bounded polytope
      for (int j = 0; j < d; ++j) z[j] += tb * p[j];
      for (int i = 0; i < n; ++i) xv[i] += tb * Np[i];
      zero_step = tb == 0.0;
      const int side = Np[jb] > 0.0 ? 1 : -1;
      if (state[jb] == 0)
        state[jb] = add_row(A, Nr + static_cast<size_t>(jb) * d, jb, side) ? 1 : 2;
      else  // a nearly dependent bound that blocks: its row is not orthogonal to p
        state[jb] = add_row(A, Nr + static_cast<size_t>(jb) * d, jb, side, 1e-12) ? 1 : 3;
      if (it % 64 == 63) eval(z, xv);
      continue;
    }
    // stationary on the face: multipliers of the active rows, R w = Q^T (-g)
    for (int l = 0; l < k; ++l) w[l] = -dot(A.Q.data() + static_cast<size_t>(l) * d, g.data(), d);
    for (int l = k - 1; l >= 0; --l) {
      double s = w[l];
      for (int q = l + 1; q < k; ++q) s -= A.R[static_cast<size_t>(q) * d + l] * w[q];
      w[l] = s / A.R[static_cast<size_t>(l) * d + l];
    }
    int drop = -1;
    double worst = -1e-10 * std::max(gn, 1e-300);
    for (int l = 0; l < k; ++l) {
      if (A.side[l] == 0) continue;
      const double mu = A.side[l] * w[l];
      if (zero_step) {
        if (mu < -1e-10 * gn && (drop < 0 || A.idx[l] < A.idx[drop])) drop = l;
      } else if (mu < worst) {
        worst = mu;
        drop = l;
      }
    }
    if (drop < 0) break;  // KKT
    state[A.idx[drop]] = 0;
    drop_row(A, drop);
    // bounds in or near the span of the rows before are candidates again
    for (int i = 0; i < n; ++i)
      if (state[i] >= 2) state[i] = 0;
  }
  // onto the active bounds exactly: z += Q R^{-T} (b - A z), kept only if
  // it is a rounding-sized move that leaves every bound met
  {
    std::vector<double> z0 = z;
    const int k = A.k();
    std::vector<double> res(k), u(k);
    for (int l = 0; l < k; ++l) {
      const int i = A.idx[l];
      const double bnd = A.side[l] >= 0 ? ub[i] : lb[i];
      res[l] = bnd - dot(Nr + static_cast<size_t>(i) * d, z.data(), d);
    }
    for (int l = 0; l < k; ++l) {  // R^T u = res (forward substitution)
      double s = res[l];
      for (int q = 0; q < l; ++q) s -= A.R[static_cast<size_t>(l) * d + q] * u[q];
      u[l] = s / A.R[static_cast<size_t>(l) * d + l];
    }
    for (int l = 0; l < k; ++l) {
      const double* ql = A.Q.data() + static_cast<size_t>(l) * d;
      for (int j = 0; j < d; ++j) z[j] += ql[j] * u[l];
    }
    eval(z, xv);
    bool ok = true;
    for (int i = 0; i < n && ok; ++i) ok = xv[i] <= ub[i] + tol_hi(i) && xv[i] >= lb[i] - tol_lo(i);
    if (!ok) z.swap(z0);
  }
  eval(z, xv);
  for (int i = 0; i < n; ++i) x[i] = std::min(std::max(xv[i], lb[i]), ub[i]);
  return it < max_iters ? 0 : 1;
}

void polish_many(int count, int threads, const std::function<void(int)>& job) {
  threads = std::max(1, std::min(threads, count));
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t] {
      for (int q = t; q < count; q += threads) job(q);
    });
  for (auto& th : pool) th.join();
}

}  // namespace achr_host

extern "C" int achr_lp_polish(int n, int d, const double* N, const double* lb, const double* ub, const double* c,
                              const double* xfeas, double* x, int max_iters) {
  if (n <= 0 || d < 0 || d > n || (d > 0 && !N) || !lb || !ub || !c || !xfeas || !x || max_iters < 0)
    return -1;
  try {
    std::vector<double> Nr(static_cast<size_t>(n) * std::max(d, 1));
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < n; ++i) Nr[static_cast<size_t>(i) * d + j] = N[static_cast<size_t>(j) * n + i];
    return achr_host::polish_lp(n, d, Nr.data(), lb, ub, c, xfeas, x, max_iters);
  } catch (const std::bad_alloc&) {
    return -7;
  }
}
